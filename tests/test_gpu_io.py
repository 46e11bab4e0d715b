"""Asynchronous trajectory output (mdpd_write_xyz_async) in the reference's
xyz format (io_util::WriteXYZForm + FPDPD::writeAscii, io_util.hpp:25-31,
ptcl_class.hpp:67-73).

Bar: the file is byte-identical to the same frames formatted by
implicit_dpd_amd.fpdpd.write_xyz (Python's correctly rounded %.15g, as
glibc's printf) from downloads of the same states, frames appended in order
while the run goes on between them; errors of the writer thread reach the
caller."""
import numpy as np
import pytest

from implicit_dpd_amd import configs
from implicit_dpd_amd.capi import MDPD_ERR_ARG, MdpdError
from implicit_dpd_amd.fpdpd import read_xyz, write_xyz

pytestmark = pytest.mark.gpu


def _engine(engine_mod, seed=5):
    box = (9.0, 8.0, 10.0)
    p, _ = configs.membrane(n_chain=1, box=box)
    s = configs.random_chains(600, box, seed, n_solvent=400)
    e = engine_mod.Engine(p, capacity=len(s["pos"]) + 64)
    e.upload(s["pos"], s["vel"], s["ids"], s["prop"], s["amp_id"], s["unit"])
    e.set_topology(s["topo"])
    e.set_clock(1, p.pair_seed)
    return e


def test_async_frames_match_reference_format(engine_mod, tmp_path):
    e = _engine(engine_mod)
    traj = tmp_path / "traject.xyz"
    want = tmp_path / "want.xyz"
    texts = []
    for k, (steps, append) in enumerate(((3, False), (2, True), (4, True))):
        e.run(steps)
        t = e.get_clock()
        rec = e.download_fpdpd()
        e.write_xyz_async(str(traj), t, append=append)
        write_xyz(str(want), rec, t)
        texts.append(want.read_text())
        # the writer runs while more steps are queued
    e.run(3)
    e.io_wait()
    got = traj.read_text()
    assert got == "".join(texts)
    # one frame read back (first frame)
    n = int(got.split("\n", 1)[0])
    first = "".join(got.splitlines(keepends=True)[: n + 2])
    one = tmp_path / "one.xyz"
    one.write_text(first)
    r, t = read_xyz(str(one))
    assert t == 4 and len(r) == n
    e.close()


def test_async_write_error_reaches_caller(engine_mod, tmp_path):
    e = _engine(engine_mod, seed=6)
    e.run(1)
    e.write_xyz_async(str(tmp_path / "no_such_dir" / "traject.xyz"), 1)
    with pytest.raises(MdpdError) as ei:
        e.io_wait()
    assert ei.value.code == MDPD_ERR_ARG
    # the writer is usable again
    e.write_xyz_async(str(tmp_path / "ok.xyz"), 1)
    e.io_wait()
    assert (tmp_path / "ok.xyz").read_text().startswith(f"{e.num_beads()}\ntime 1\n")
    e.close()
