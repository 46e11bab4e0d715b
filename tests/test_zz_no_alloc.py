"""mdpd_rdf_noalloc (the allocation-free variant of mdpd_rdf, its scratch in
one fixed device array per GPU) against mdpd_rdf through the C ABI: the
same counts bit for bit, on two calls in a row where the second has the
smaller input (fewer beads and bins, one replica), so any counter the first
call left behind would show; then the size limit of the fixed scratch, which
must be refused before any work."""
import ctypes as C

import numpy as np
import pytest

from implicit_dpd_amd import capi, configs

pytestmark = pytest.mark.gpu


def _rdf_both(e, rmax, nbins):
    L = capi.lib()
    a = np.full((e.n_rep, nbins), -1.0)
    b = np.full((e.n_rep, nbins), -2.0)
    assert L.mdpd_rdf(e.h, rmax, nbins, a.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert L.mdpd_rdf_noalloc(e.h, rmax, nbins, b.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return a, b


def test_rdf_noalloc_equals_rdf_twice_then_smaller(engine_mod):
    # call 1: a 4-replica ensemble, 300 bins
    sizes = [3000, 5000, 4000, 2000]
    systems, params = [], []
    for r, nb in enumerate(sizes):
        box = (9.0, 9.0, 8.0 + r)
        params.append(configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=10 + r))
        systems.append(configs.random_fluid(nb, box, 20 + r))
    s = configs.concat_systems(systems)
    e = engine_mod.Engine(params, n_beads=sizes)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_clock(0, 10)
    e.run(5)
    a, b = _rdf_both(e, 1.0, 300)
    assert a.sum() > 0 and np.array_equal(a, b)
    # call 2: a smaller single system with fewer bins, right after
    box = (7.0, 7.0, 7.0)
    p = configs.warren_params(box, A=-40.0, B=25.0, rd=0.75, pair_seed=3)
    f = configs.random_fluid(1500, box, 7)
    e2 = engine_mod.Engine(p, capacity=1500)
    e2.upload(f["pos"], f["vel"], f["ids"], f["prop"], f["amp_id"], f["unit"])
    e2.set_clock(0, 3)
    e2.run(3)
    a2, b2 = _rdf_both(e2, 0.8, 40)
    assert a2.sum() > 0 and np.array_equal(a2, b2)
    # and the first context again, still equal
    a3, b3 = _rdf_both(e, 1.0, 300)
    assert np.array_equal(a3, b3) and np.array_equal(a3, a)
    e.close()
    e2.close()


def test_rdf_noalloc_refuses_oversized_requests(engine_mod):
    sizes = [2000] * 80
    params = [configs.warren_params((7.0, 7.0, 7.0), A=-40.0, B=25.0, rd=0.75, pair_seed=1 + r)
              for r in range(len(sizes))]
    s = configs.concat_systems([configs.random_fluid(2000, (7.0, 7.0, 7.0), 40 + r)
                                for r in range(len(sizes))])
    e = engine_mod.Engine(params, n_beads=sizes)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_clock(0, 1)
    e.run(2)
    L = capi.lib()
    out = np.zeros((80, 1024))
    # 80 * 1024 counters > 65536: refused before any work, output untouched
    rc = L.mdpd_rdf_noalloc(e.h, 1.0, 1024, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == capi.MDPD_ERR_CAPACITY and not out.any()
    assert b"65536" in L.mdpd_last_error(e.h)
    # 80 * 800 = 64000 fits, and matches mdpd_rdf
    a, b = _rdf_both(e, 1.0, 800)
    assert np.array_equal(a, b) and a.sum() > 0
    rc = L.mdpd_rdf_noalloc(e.h, 1.0, 0, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == capi.MDPD_ERR_ARG
    e.close()
