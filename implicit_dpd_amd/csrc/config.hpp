// Initial-configuration generator shared by the host code (config.cpp) and
// the kernels (k_config.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mdpd_capi.h"

namespace mdpd {

// keys of the per-object Saru streams (seed2 of Saru(index, tag, seed))
constexpr uint32_t kTagAmp = 0x414D5048u;  // amphiphile base point and direction
constexpr uint32_t kTagSol = 0x534F4C56u;  // solvent position attempts
constexpr uint32_t kTagVel = 0x56454C4Fu;  // velocities

// a solvent bead gives up after this many rejected positions (the reference
// would loop for ever when its region is empty)
constexpr uint32_t kSolMaxAttempts = 1u << 16;

enum : uint32_t { kCfgFlagOutOfBox = 1u, kCfgFlagSolvent = 2u };
enum : int32_t { kSolNone = 0, kSolCuboid = 1, kSolSphere = 2, kSolCylinder = 3 };

// One ring of a sphere or cylinder leaflet (MakeSphLine): amphiphiles
// start .. start + count - 1 at phi = d_phi * k, k = 0 .. count - 1.
struct CfgRing {
  double sin_t, cos_t;  // sine and cosine of the polar angle
  double rad, sign;     // leaflet radius argument and orientation (-1 outer, +1 inner)
  double d_phi;         // 2 pi / phi_elem
  double cz;            // z of the ring's centre
  uint32_t start, count;
};

// where generated beads go: pos/vel component c of bead i at
// [i * bead_stride + c * comp_stride]
struct CfgOut {
  double* pos;
  double* vel;
  double* vbuf;  // vel_buf (may be null)
  uint32_t* id;
  uint32_t* prop;
  uint32_t* amp_id;
  uint32_t* unit;
  size_t bead_stride, comp_stride;
};

struct CfgLaunch {
  int32_t mode;
  uint32_t n, amp_num, all_unit, head_unit;
  uint32_t rat;          // ASYM_FLAT: (uint)(100 in_out_rat)
  uint32_t seed;
  double box[3];
  double center[2];      // x, y of sphere / cylinder centres
  double bond_leng;
  double span;           // FLAT: (all_unit - 1) bond_leng + 1e-5
  double q_thick;        // 0.5 all_unit bond_leng
  double sd;             // sqrt(Temperature)
  const CfgRing* rings;
  uint32_t n_rings;
  int32_t sol_kind, sol_axis;
  double sol_lo[3], sol_hi[3];  // cuboid solvent region
  double sol_rad;               // spherical solvent: accept dist < sol_rad
  uint32_t* flags;
  CfgOut out;
};

// the beads, the centre-of-mass drift removal and vel_buf = vel; scratch
// holds cfg_scratch_doubles(n) doubles, and afterwards [0..2] = sum of v
// before the removal, [3] = sum of v.v after it
void launch_cfg_generate(const CfgLaunch& L, double* scratch, hipStream_t st);
inline size_t cfg_scratch_doubles(uint32_t n) { return 8 + 6 * (size_t)((n + 255) / 256); }
void launch_cfg_chains(uint32_t n, uint32_t amp_num, double cf_b, uint32_t* chain, uint32_t* orig,
                       uint32_t* topo, double* cf_bend, uint32_t* amp_rep, hipStream_t st);

}  // namespace mdpd
