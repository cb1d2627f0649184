// Grouped launch of the fused FP32 pair loops: one kernel launch per df stage runs the
// quadrature-pair loops of every fused grid shape of an uneven batch (3D 32^3 real-data
// form, 3D 16^3, 2D 64^2, 2D 32^2 and 2D 16^2 C2C forms).
//
// The launch is a list of segments, one per plan.  Every workgroup has 512 threads and the
// LDS of the largest body (the 3D 32^3 form); a segment of a smaller body runs VBS =
// 512 / T of its T-thread workgroups ("virtual blocks") side by side in one workgroup, each
// with its own slice of the LDS.  The virtual blocks of one workgroup share its barriers,
// so they run the same number of pairs: they take the same angle range (k_fusedc3,
// k_fusedw) or the same round count (k_fused2).  A workgroup whose last virtual blocks
// have no unit of work repeats the segment's last unit: the repeat writes the same values
// to the same partial, so the result is unchanged.  Segments start at multiples of 8
// workgroups, so the XCD numbering of the 3D 32^3 form (cell_of_block) holds inside the
// launch.
#pragma once
#include "fused2_body.h"
#include "fused3_body.h"

namespace ib {

enum GroupKind { GK_F3 = 0, GK_F3K, GK_C3, GK_W64, GK_W64S, GK_F2_16, GK_F2_32 };

constexpr int kGroupThreads = 512;
constexpr int kMaxGroupSegs = 8;

struct GroupSeg {
  int kind;
  long long blocks;  // workgroups of the segment (a multiple of 8)
  long long units;   // virtual blocks with work: (cell, j) or (cell, w) pairs
  union {
    F3Args f3;
    C3Args c3;
    CWArgs<float> w;
    F2Args f2;
  };
};

struct GroupArgs {
  int nseg;
  GroupSeg seg[kMaxGroupSegs];
};

// threads of one virtual block of a kind
__host__ __device__ constexpr int group_vb_threads(int kind) {
  return kind == GK_F3 || kind == GK_F3K ? 512
         : kind == GK_C3                 ? c3::THREADS
         : kind == GK_W64 || kind == GK_W64S ? 128
         : kind == GK_F2_16              ? F2Shape<16>::THREADS
                                         : F2Shape<32>::THREADS;
}

// workgroups of a segment: angle-range-major for the split C2C kinds (all virtual blocks
// of a workgroup share the angle range), rounded up to a multiple of 8
inline long long group_blocks(int kind, long long units, int ns, long long f3_grid) {
  long long b;
  if (kind == GK_F3 || kind == GK_F3K) {
    b = f3_grid;
  } else {
    const int vbs = kGroupThreads / group_vb_threads(kind);
    b = (units + vbs - 1) / vbs * (kind == GK_F2_16 || kind == GK_F2_32 ? 1 : ns);
  }
  return (b + 7) / 8 * 8;
}

// Launch the grouped kernel over n segments (collide_group.hip).
int launch_fused_group(const GroupSeg* segs, int n, hipStream_t st);

}  // namespace ib
