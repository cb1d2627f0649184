// The grouped launch of the fused FP32 pair loops (layout and rules in group.h): one
// kernel, k_fused_group, whose workgroups run the body of their segment's grid shape.
#include "group.h"
#include "plan.h"

namespace ib {

namespace gl {
// LDS layout: byte offsets of each body's arrays inside the kernel's LDS bytes, computed
// from the array types so that they match the bodies' parameters exactly.
template <size_t A>
constexpr size_t up16() {
  return (A + 15) / 16 * 16;
}
// 3D 32^3 real-data form (one body per workgroup)
using F3W = cx<float>[f3::NPL][f3::N][f3::PAD];
using F3T = cx<float>[4][3][f3::N];
using F3Sp = float[3 * (f3::N / 2) * (f3::N / 2) + 1];
using F3P = cx<float>[f3::N][f3::PAD];
constexpr size_t F3_TE = up16<sizeof(F3W)>();
constexpr size_t F3_TG = F3_TE + up16<sizeof(F3T)>();
constexpr size_t F3_SP = F3_TG + up16<sizeof(F3T)>();
constexpr size_t F3_S16 = F3_SP + up16<sizeof(F3Sp)>();
constexpr size_t F3_SPL = F3_S16 + up16<sizeof(F3P)>();
constexpr size_t F3_END = F3_SPL + up16<sizeof(F3P)>();
// 3D 16^3 C2C form, per virtual block
using C3W = cx<float>[c3::N][c3::N][c3::PAD];
using C3T = cx<float>[2][3][c3::N];
using C3Sp = float[3 * (c3::N / 2) * (c3::N / 2) + 1];
constexpr size_t C3_TE = up16<sizeof(C3W)>();
constexpr size_t C3_TG = C3_TE + up16<sizeof(C3T)>();
constexpr size_t C3_SP = C3_TG + up16<sizeof(C3T)>();
constexpr size_t C3_END = C3_SP + up16<sizeof(C3Sp)>();
// 2D 64^2 two-lane form, per virtual block
using WW = cx<float>[64 * 64];
using WT = cx<float>[2][2][64];
using WTw = cx<float>[2][64 / 2 + 1];
constexpr size_t W_TE = up16<sizeof(WW)>();
constexpr size_t W_TG = W_TE + up16<sizeof(WT)>();
constexpr size_t W_TW = W_TG + up16<sizeof(WT)>();
constexpr size_t W_END = W_TW + up16<sizeof(WTw)>();
// 2D 16^2 / 32^2 form, per virtual block
template <int N>
struct F2Off {
  static constexpr size_t TAB = up16<sizeof(typename F2L<N>::Wg)>();
  static constexpr size_t FHS = TAB + up16<sizeof(typename F2L<N>::Tab)>();
  static constexpr size_t FS = FHS + up16<sizeof(typename F2L<N>::Fhs)>();
  static constexpr size_t END = FS + up16<sizeof(typename F2L<N>::Fs)>();
};

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t BYTES =
    cmax(cmax(F3_END, 2 * C3_END), cmax(4 * W_END, cmax(2 * F2Off<16>::END, 4 * F2Off<32>::END)));
static_assert(BYTES <= 160 * 1024, "grouped fused kernel LDS exceeds 160 KiB");
static_assert(kGroupThreads / c3::THREADS == 2 && kGroupThreads / F2Shape<32>::THREADS == 4 &&
                  kGroupThreads / F2Shape<16>::THREADS == 2,
              "virtual block counts assumed by the LDS layout");

template <typename A>
__device__ __forceinline__ A& at(char* base, size_t off) {
  return *reinterpret_cast<A*>(base + off);
}
}  // namespace gl

// (cell, partial index) of virtual block vb of workgroup b of a split C2C segment: all
// virtual blocks of a workgroup take angle range b mod ns; the unit index is clamped to
// the segment's last unit (see group.h)
__device__ __forceinline__ void group_unit(long long b, int vb, int vbs, int ns, int nr, long long units,
                                          long long& cell, int& jj) {
  const int asp = (int)(b % ns);
  long long u = (b / ns) * vbs + vb;
  if (u >= units) u = units - 1;
  cell = u / nr;
  jj = (int)(u % nr) * ns + asp;
}

__global__ __launch_bounds__(kGroupThreads, 1) void k_fused_group(GroupArgs G) {
  __shared__ __align__(16) char lds[gl::BYTES];
  long long b = blockIdx.x;
  int s = 0;
  while (s + 1 < G.nseg && b >= G.seg[s].blocks) {  // uniform over the workgroup
    b -= G.seg[s].blocks;
    ++s;
  }
  const GroupSeg& S = G.seg[s];
  const int tid = threadIdx.x;
  switch (S.kind) {
    case GK_F3:
    case GK_F3K: {
      long long cell;
      int jj;
      cell_of_block(b, S.f3.nr * S.f3.ns, S.f3.xg, cell, jj);
      if (cell >= S.f3.batch) return;  // uniform over the workgroup
      using namespace gl;
      if (S.kind == GK_F3K)
        fused3_body<true>(S.f3, cell, jj, tid, at<F3W>(lds, 0), at<F3T>(lds, F3_TE), at<F3T>(lds, F3_TG),
                          at<F3Sp>(lds, F3_SP), at<F3P>(lds, F3_S16), at<F3P>(lds, F3_SPL));
      else
        fused3_body<false>(S.f3, cell, jj, tid, at<F3W>(lds, 0), at<F3T>(lds, F3_TE), at<F3T>(lds, F3_TG),
                           at<F3Sp>(lds, F3_SP), at<F3P>(lds, F3_S16), at<F3P>(lds, F3_SPL));
      return;
    }
    case GK_C3: {
      constexpr int T = c3::THREADS, VBS = kGroupThreads / T;
      if ((b / S.c3.ns) * VBS >= S.units) return;  // padding workgroup
      const int vb = __builtin_amdgcn_readfirstlane(tid / T);
      long long cell;
      int jj;
      group_unit(b, vb, VBS, S.c3.ns, S.c3.nr, S.units, cell, jj);
      using namespace gl;
      char* base = lds + vb * C3_END;
      fusedc3_body(S.c3, cell, jj, tid - vb * T, at<C3W>(base, 0), at<C3T>(base, C3_TE), at<C3T>(base, C3_TG),
                   at<C3Sp>(base, C3_SP));
      return;
    }
    case GK_W64:
    case GK_W64S: {
      constexpr int T = 128, VBS = kGroupThreads / T;
      if ((b / S.w.ns) * VBS >= S.units) return;  // padding workgroup
      const int vb = __builtin_amdgcn_readfirstlane(tid / T);
      long long cell;
      int jj;
      group_unit(b, vb, VBS, S.w.ns, S.w.nr, S.units, cell, jj);
      using namespace gl;
      char* base = lds + vb * W_END;
      // the general (angle-range) body serves both kinds: one instance of the 2D 64^2 body in
      // this kernel instead of two
      fusedw_body<float, 64, true>(S.w, cell, jj, tid - vb * T, at<WW>(base, 0), at<WT>(base, W_TE),
                                   at<WT>(base, W_TG), at<WTw>(base, W_TW));
      return;
    }
    case GK_F2_16:
    case GK_F2_32: {
      const bool n16 = S.kind == GK_F2_16;
      const int T = n16 ? F2Shape<16>::THREADS : F2Shape<32>::THREADS, VBS = kGroupThreads / T;
      if (b * VBS >= S.units) return;  // padding workgroup
      const int vb = __builtin_amdgcn_readfirstlane(tid / T);
      long long u = b * VBS + vb;
      if (u >= S.units) u = S.units - 1;
      const long long cell = u / S.f2.wpc;
      const int w = (int)(u % S.f2.wpc);
      using namespace gl;
      if (n16) {
        using O = F2Off<16>;
        char* base = lds + vb * O::END;
        fused2_body<16>(S.f2, cell, w, tid - vb * T, at<F2L<16>::Wg>(base, 0), at<F2L<16>::Tab>(base, O::TAB),
                        at<F2L<16>::Fhs>(base, O::FHS), at<F2L<16>::Fs>(base, O::FS));
      } else {
        using O = F2Off<32>;
        char* base = lds + vb * O::END;
        fused2_body<32>(S.f2, cell, w, tid - vb * T, at<F2L<32>::Wg>(base, 0), at<F2L<32>::Tab>(base, O::TAB),
                        at<F2L<32>::Fhs>(base, O::FHS), at<F2L<32>::Fs>(base, O::FS));
      }
      return;
    }
    default:
      return;
  }
}

int launch_fused_group(const GroupSeg* segs, int n, hipStream_t st) {
  if (n < 1 || n > kMaxGroupSegs) {
    set_error("grouped fused launch: 1 to 8 segments");
    return IB_ERR_ARG;
  }
  GroupArgs G{};
  G.nseg = n;
  long long total = 0;
  for (int i = 0; i < n; ++i) {
    G.seg[i] = segs[i];
    total += segs[i].blocks;
  }
  if (total <= 0) return IB_OK;
  if (total > 0x7fffffffLL) {
    set_error("grouped fused launch: grid too large");
    return IB_ERR_ARG;
  }
  hipLaunchKernelGGL(k_fused_group, dim3((unsigned)total), dim3(kGroupThreads), 0, st, G);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
