// rhs_pencil.h — the RHS of one face direction per launch, with the stencil
// staged through LDS in batches of variables.
//
// A workgroup owns a tile of TA x TB x TC faces (TA along the face direction
// "a", TB/TC along the two other directions "b"/"c" in increasing axis
// order).  Consecutive faces along a share a cell, so TA faces difference
// into TA-1 cells and neighbouring tiles overlap by one face.
//
// Inviscid phase (InviscidScheme.cu:24-108 via Reconstruction.cu:12-89):
// the 5+NS reconstructed variables are staged over the box (TA+3) x TB x TC
// (the 4-point MUSCL stencil along a), KI variables per batch.  Batches go
// straight into two alternating LDS buffers with global-to-LDS loads, the
// next batch in flight while the current one is used (one barrier per
// batch).
// Viscous phase (ViscousScheme.cu:11-342): u, v, w, T and the mass fractions
// are staged over (TA+1) x (TB+2) x (TC+2) (the face points plus the
// tangential neighbours of the 4-point averages), KV per batch.  Face-point
// data (metric, Jacobian, mu, kappa, rho*D) is read directly with coalesced
// loads: threads are laid out x-fastest for every direction.  Each thread
// keeps its face flux F_inv - F_visc in an LDS slot; the cell difference is
// taken once at the end and dq written (first direction) or accumulated.
#pragma once

#include <utility>

#include "device_physics.h"

namespace cfx {

// f(integral_constant<int, I>) for I = 0..N-1, unrolled at compile time.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Two consecutive elements of a field (the two points of an x face); element
// alignment only, so the load may straddle a 16-byte boundary.
template <typename R>
struct alignas(sizeof(R)) FacePair {
  R l, r;
};

template <int DIR>
struct Axes;
template <>
struct Axes<0> {
  static constexpr int a = 0, b = 1, c = 2;
};
template <>
struct Axes<1> {
  static constexpr int a = 1, b = 0, c = 2;
};
template <>
struct Axes<2> {
  static constexpr int a = 2, b = 0, c = 1;
};

// A staging box: extents (Ea, Eb, Ec) along the face axes a, b, c, laid out
// x-fastest in LDS.
template <int DIR, int Ea, int Eb, int Ec>
struct Box {
  static constexpr int A = Axes<DIR>::a, Bx = Axes<DIR>::b, Cx = Axes<DIR>::c;
  static constexpr int L0 = A == 0 ? Ea : (Bx == 0 ? Eb : Ec);
  static constexpr int L1 = A == 1 ? Ea : (Bx == 1 ? Eb : Ec);
  static constexpr int L2 = A == 2 ? Ea : (Bx == 2 ? Eb : Ec);
  static constexpr int V = L0 * L1 * L2;
  static constexpr int str[3] = {1, L0, L0 * L1};
  static constexpr int sa = str[A], sb = str[Bx], sc = str[Cx];
};

template <int NS, int DIR, int TA, int TB, int TC, typename R, int WPE>
struct PencilShape {
  static constexpr int A = Axes<DIR>::a, Bx = Axes<DIR>::b, Cx = Axes<DIR>::c;
  static constexpr int NV = 5 + NS;
  static constexpr int NT = TA * TB * TC;
  static constexpr int T0 = A == 0 ? TA : (Bx == 0 ? TB : TC);
  static constexpr int T1 = A == 1 ? TA : (Bx == 1 ? TB : TC);
  static constexpr int T2 = A == 2 ? TA : (Bx == 2 ? TB : TC);
  static constexpr int tx[3] = {1, T0, T0 * T1};
  using BI = Box<DIR, TA + 3, TB, TC>;
  using BV = Box<DIR, TA + 1, TB + 2, TC + 2>;
  static constexpr bool F32 = sizeof(R) == 4;
  // The tiles stage through two LDS buffers filled by global-to-LDS loads
  // (no registers in between): the next batch is in flight while the
  // current one is used.  A wavefront's dword load writes 64 consecutive
  // dwords, EPW elements (FP64: two lanes per element, one per word), so
  // every variable's slice is padded to whole wavefront loads.
  static constexpr int W = (int)sizeof(R) / 4;  // dwords per element
  static constexpr int EPW = 64 / W;
  static constexpr int VPI = (BI::V + EPW - 1) / EPW * EPW;  // LDS stride of one inviscid variable
  static constexpr int VPV = (BV::V + EPW - 1) / EPW * EPW;  // ... of one viscous variable
  // batch sizes: 7 inviscid variables; 3 viscous, 2 where the FP64 buffers
  // would pass the LDS share of the tile's occupancy target (measured:
  // batches that fill the share exactly cost a workgroup per CU)
  static constexpr int KI = NV < 7 ? NV : 7;
  static constexpr int lds_share = 163840 / (WPE * 256 / NT > 0 ? WPE * 256 / NT : 1) - 2048;
  static constexpr int lds_bytes(int kv) {
    return (2 * (KI * VPI > kv * VPV ? KI * VPI : kv * VPV) + NV * NT) * (int)sizeof(R);
  }
  static constexpr int KV = F32 || lds_bytes(3) <= lds_share ? 3 : 2;
  static constexpr int STG = KI * VPI > KV * VPV ? KI * VPI : KV * VPV;
};

// Tile order inside a block: the tiles of the (b, c) plane are visited in
// G x G groups with a fastest, so that tiles sharing a tangential halo run
// close together; G = 0 visits c fastest, then a, then b (neighbours along c
// share two of the three rows of the viscous box).  Returns false for the
// padding tiles of partial groups.
template <int G>
__device__ __forceinline__ bool pencil_tile(int lid, int nta, int ntb, int ntc, int &ia, int &ib, int &ic) {
  if constexpr (G == 0) {
    ic = lid % ntc;
    const int q = lid / ntc;
    ia = q % nta;
    ib = q / nta;
    return true;
  }
  const int gb = (ntb + G - 1) / G;
  ia = lid % nta;
  int q = lid / nta;
  const int bin = q % G;
  q /= G;
  const int cin = q % G;
  q /= G;
  ib = (q % gb) * G + bin;
  ic = (q / gb) * G + cin;
  return ib < ntb && ic < ntc;
}

// One tile of the RHS of face direction DIR: tseq is the tile's position in
// the batch's tile sequence (block by block, off[] the per-block prefix).
// Workgroups are dealt to the 8 XCDs round-robin, so position t is remapped to
// lid = (t % 8) * (padded / 8) + t / 8: each XCD works through a contiguous
// run of tiles (whole blocks where they are small) and keeps their halos in
// its own L2.
template <int NS, int DIR, int TA, int TB, int TC, typename R, int G, int WPE>
__device__ __forceinline__ void rhs_pencil_tile(const BlockView *__restrict__ views, const int *__restrict__ off,
                                                int nb, int padded, int tseq, SpeciesPtrT<R> sp, const SchemeDev &S,
                                                int accumulate, R *__restrict__ stg0, R *__restrict__ stg1,
                                                R *__restrict__ fbuf) {
  using P = PencilShape<NS, DIR, TA, TB, TC, R, WPE>;
  using BI = typename P::BI;
  using BV = typename P::BV;
  constexpr int NV = P::NV, NT = P::NT, A = P::A, Bx = P::Bx, Cx = P::Cx;
  constexpr int ta_stride = P::tx[A];
  using Fd = Fields<R>;
  const int total = off[nb] - off[0];
  {
    const int lid = (tseq & 7) * (padded >> 3) + (tseq >> 3);
    if (lid >= total) return;  // uniform per workgroup
    // uniform per workgroup: keep the block index and the view in SGPRs
    const int wgi = __builtin_amdgcn_readfirstlane(lid + off[0]);
    const int blk = __builtin_amdgcn_readfirstlane(find_seg(off, nb, wgi));
    const GView B = gview(views[blk]);
    const gptr<const R> __restrict__ BV_ = Fd::bv(B);
    const gptr<const R> __restrict__ YK_ = Fd::yk(B);
    const gptr<const R> __restrict__ PR_ = Fd::props(B);
    const gptr<const R> __restrict__ MT_ = Fd::metric(B);
    const gptr<const R> __restrict__ JC_ = Fd::jac(B);

    const long long np = B.np;
    const int m[3] = {B.mx, B.my, B.mz};
    const long long st[3] = {1, B.sy, B.sz};
    const int nta = (m[A] + TA - 2) / (TA - 1);  // TA faces cover TA-1 cells
    const int ntb = (m[Bx] + TB - 1) / TB;
    const int ntc = (m[Cx] + TC - 1) / TC;
    int ia, ib, ic;
    if (!pencil_tile<G>(wgi - off[blk], nta, ntb, ntc, ia, ib, ic)) return;  // uniform per workgroup
    int o[3];
    o[A] = ia * (TA - 1);
    o[Bx] = ib * TB;
    o[Cx] = ic * TC;
    const int tid = threadIdx.x;
    int t[3];
    t[0] = tid % P::T0;
    t[1] = (tid / P::T0) % P::T1;
    t[2] = tid / (P::T0 * P::T1);
    int g[3];
    g[A] = o[A] - 1 + t[A];
    g[Bx] = o[Bx] + t[Bx];
    g[Cx] = o[Cx] + t[Cx];
    const bool face_ok = g[A] + 1 <= m[A] && g[Bx] < m[Bx] && g[Cx] < m[Cx];
    const bool cell_ok = face_ok && t[A] <= TA - 2 && g[A] + 1 < m[A];
    const long long pl_i = face_ok ? (long long)(g[2] + kNgg) * B.sz + (long long)(g[1] + kNgg) * B.sy + (g[0] + kNgg) : 0;
    const long long pr_i = pl_i + st[A];
    // byte offsets of the two face points (32-bit: shared by every field array)
    const unsigned pl_b = (unsigned)pl_i * (unsigned)sizeof(R), pr_b = (unsigned)pr_i * (unsigned)sizeof(R);
    // a field at the face's two points: for x faces they are adjacent and
    // come in one load of two elements
    auto ld2 = [&](gptr<const R> base, R &l, R &r) {
      if constexpr (DIR == 0) {
        const gptr<const FacePair<R>> v = (gptr<const FacePair<R>>)((gptr<const char>)base + pl_b);
        l = v->l, r = v->r;
      } else {
        l = ldg<R>(base, pl_b), r = ldg<R>(base, pr_b);
      }
    };
    const int szi = (int)B.sz, syi = (int)B.sy;

    R k1 = 0, k2 = 0, k3 = 0;
    // face averages of the metric row of the face direction, formed with k
    // from the same loads and kept for the viscous flux
    R mfa[3] = {0, 0, 0};
    const bool q5 = (S.quirks & CFX_Q_METRIC_AVG) && DIR != 1;
    // ---- batches through two LDS buffers.  Batch n lands in buffer n % 2;
    // after the barrier that publishes batch n, batch n + 1 is issued into
    // the other buffer (whose last readers passed that barrier) and stays
    // in flight while batch n is used.  Both buffers are free when a tile
    // starts, so the first two batches go out together.  The buffers are
    // separate shared arrays: the compiler's waits for global-to-LDS loads
    // then tell them apart and do not hold the reads of one buffer until
    // the loads into the other have landed.
    constexpr int VPI = P::VPI, VPV = P::VPV, KI = P::KI, KV = P::KV;
    constexpr int NI = (NV + KI - 1) / KI;         // inviscid batches
    constexpr int NVV = 4 + NS;                    // viscous variables: u, v, w, T, Y_l
    constexpr int NVB = (NVV + KV - 1) / KV;       // viscous batches
    const bool viscous = S.viscous == 2;
    R *const buf0 = stg0, *const buf1 = stg1;
    // The byte offsets of this lane's staged elements: lane e of a box
    // loads element e of each variable, the wavefront writing 64 consecutive
    // slots from its first element on (FP64: lane pairs share an element,
    // one loading each word).  Computed once per tile for each box and kept
    // for all its batches.
    constexpr int W = P::W, EPW = P::EPW, NW = NT / 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    // the range readfirstlane hides: the chunks of whole rounds of NW are
    // then issued without a test
    __builtin_assume(wave >= 0 && wave < NW);
    auto offsets = [&](auto box, const int bo[3], unsigned *gb) {
      using BX = decltype(box);
      constexpr int NC = (BX::V + EPW - 1) / EPW;  // wavefront loads per variable
      constexpr int IT = (NC + NW - 1) / NW;
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int e0 = (it * NW + wave) * EPW + lane / W;
        const int e = e0 < BX::V ? e0 : BX::V - 1;
        const int x = min(bo[0] + e % BX::L0, m[0] + 1);
        const int y = min(bo[1] + (e / BX::L0) % BX::L1, m[1] + 1);
        const int z = min(bo[2] + e / (BX::L0 * BX::L1), m[2] + 1);
        gb[it] = (unsigned)((z + kNgg) * szi + (y + kNgg) * syi + (x + kNgg)) * (unsigned)sizeof(R) +
                 (unsigned)(lane % W) * 4u;
      }
    };
    // issue variables [v0, v0 + K) of box BX (LDS stride VP) into buf
    auto issue = [&](auto box, auto kc, auto vpc, auto srcf, const unsigned *gb, R *buf) {
      using BX = decltype(box);
      constexpr int K = decltype(kc)::value, VP = decltype(vpc)::value;
      constexpr int NC = (BX::V + EPW - 1) / EPW;
      constexpr int IT = (NC + NW - 1) / NW;
      // raised wavefront priority while the batch's loads are issued, so that
      // they go out ahead of other wavefronts' arithmetic (FP32 step -0.9%)
      __builtin_amdgcn_s_setprio(2);
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int chunk = it * NW + wave;
        if (chunk < NC) {  // uniform per wavefront
          // the offset enters this block as a 32-bit value, so the loads
          // use the scalar-base + 32-bit-offset form (a 64-bit offset kept
          // from the start of the tile would need a 64-bit add per load)
          unsigned go = gb[it];
          asm volatile("" : "+v"(go));
#pragma unroll
          for (int k = 0; k < K; ++k)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)((gptr<const char>)srcf(k) + go),
                (__attribute__((address_space(3))) void *)(buf + k * VP + chunk * EPW), 4, 0, 0);
        }
      }
      __builtin_amdgcn_s_setprio(0);
    };
    // the loads of this thread (and so of the batch, after the barrier)
    // have written LDS, except the last `later` ones it issued: vmcnt(later)
    // with lgkmcnt and expcnt left open.  Vector memory loads complete in
    // issue order, and the issues of different batches stay in issue order
    // (s_setprio, which brackets each batch, is a scheduling boundary).
    // With later > 0 the barrier goes without __syncthreads' workgroup
    // fences, which would wait for every global-to-LDS load in flight
    // (vmcnt(0)); the empty asm statements with a memory clobber keep the
    // compiler from moving memory accesses across it instead.  Used only
    // where no LDS store of this tile precedes the barrier.
    auto land = [&](auto later) {
      constexpr int n = decltype(later)::value;
      static_assert(n >= 0 && n < 64, "vmcnt holds 6 bits");
      __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | 0x0F70);
      if constexpr (n == 0) {
        __syncthreads();
      } else {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
      }
    };
    using none = std::integral_constant<int, 0>;
    int boi[3], bov[3];
    boi[A] = o[A] - 2, boi[Bx] = o[Bx], boi[Cx] = o[Cx];
    bov[A] = o[A] - 1, bov[Bx] = o[Bx] - 1, bov[Cx] = o[Cx] - 1;
    constexpr int ITI = ((BI::V + EPW - 1) / EPW + NW - 1) / NW, ITV = ((BV::V + EPW - 1) / EPW + NW - 1) / NW;
    unsigned gbi[ITI], gbv[ITV];
    offsets(BI{}, boi, gbi);
    if (viscous) offsets(BV{}, bov, gbv);
    auto issue_inv = [&](auto nc, R *buf) {
      constexpr int v0 = decltype(nc)::value * KI;
      constexpr int K = NV - v0 < KI ? NV - v0 : KI;
      issue(BI{}, std::integral_constant<int, K>{}, std::integral_constant<int, VPI>{},
            [&](int k) { return v0 + k < 5 ? BV_ + (v0 + k) * np : YK_ + (v0 + k - 5) * np; }, gbi, buf);
    };
    auto issue_vis = [&](auto nc, R *buf) {
      constexpr int j0 = decltype(nc)::value * KV;
      constexpr int K = NVV - j0 < KV ? NVV - j0 : KV;
      issue(BV{}, std::integral_constant<int, K>{}, std::integral_constant<int, VPV>{},
            [&](int k) { return j0 + k < 4 ? BV_ + (j0 + k < 3 ? j0 + k + 1 : 5) * np : YK_ + (j0 + k - 4) * np; },
            gbv, buf);
    };

    // ---- inviscid phase: F_inv into this thread's face slot
    {
      int l[3];
      l[A] = t[A] + 1;
      l[Bx] = t[Bx];
      l[Cx] = t[Cx];
      const int li = l[0] + BI::L0 * (l[1] + BI::L1 * l[2]);
      constexpr int sa = BI::sa;
      R pl[NV], pr[NV];
      // the face-area data (Jacobian and metric row of the face direction
      // at the face's two points), loaded ahead of the staged batches so
      // that its wait is the first batch's (lanes off the block read at
      // offset 0 and do not use it)
      R fj[2], fm[3][2];
      ld2(JC_, fj[0], fj[1]);
#pragma unroll
      for (int r = 0; r < 3; ++r) ld2(MT_ + (3 * A + r) * np, fm[r][0], fm[r][1]);
      issue_inv(std::integral_constant<int, 0>{}, buf0);
      if constexpr (NI > 1) issue_inv(std::integral_constant<int, 1>{}, buf1);
      static_for<NI>([&](auto bi) {
        constexpr int n = decltype(bi)::value;
        constexpr int v0 = n * KI;
        constexpr int K = NV - v0 < KI ? NV - v0 : KI;
        R *const cur = n % 2 == 0 ? buf0 : buf1;
        R *const nxt = n % 2 == 0 ? buf1 : buf0;
        if constexpr (n == 0 && NI > 1) {
          // batch 1 stays in flight: every wavefront issued at least
          // NC / NW of its wavefront loads for each of its variables
          constexpr int K1 = NV - KI < KI ? NV - KI : KI;
          constexpr int NC = (BI::V + EPW - 1) / EPW;
          land(std::integral_constant<int, K1 * (NC / NW)>{});
        } else {
          land(none{});
        }
        if constexpr (n + 1 == NI) {
          if (viscous) issue_vis(std::integral_constant<int, 0>{}, nxt);
        } else if constexpr (n > 0) {
          issue_inv(std::integral_constant<int, n + 1>{}, nxt);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const R *q = cur + k * VPI + li;
          muscl<R>(S, q[-sa], q[0], q[sa], q[2 * sa], pl[v0 + k], pr[v0 + k]);
        }
      });
      R F[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) F[v] = R(0);
      if (face_ok) {
        const R jl = fj[0], jr = fj[1];
        k1 = R(0.5) * (jl * fm[0][0] + jr * fm[0][1]);
        mfa[0] = (q5 && 3 * A != 0) ? R(0.5) * (fm[0][0] + fm[0][0]) : R(0.5) * (fm[0][0] + fm[0][1]);
        k2 = R(0.5) * (jl * fm[1][0] + jr * fm[1][1]);
        mfa[1] = q5 ? R(0.5) * (fm[1][0] + fm[1][0]) : R(0.5) * (fm[1][0] + fm[1][1]);
        k3 = R(0.5) * (jl * fm[2][0] + jr * fm[2][1]);
        mfa[2] = q5 ? R(0.5) * (fm[2][0] + fm[2][0]) : R(0.5) * (fm[2][0] + fm[2][1]);
        ausm_from_states<NS, R>(sp, ThermoConst<R>{sp}, S, pl, pr, k1, k2, k3, F);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) fbuf[v * NT + tid] = F[v];
    }

    // ---- viscous phase: F_visc subtracted from the slot as it is formed
    if (viscous) {
      int l[3];
      l[A] = t[A];
      l[Bx] = t[Bx] + 1;
      l[Cx] = t[Cx] + 1;
      const int li = l[0] + BV::L0 * (l[1] + BV::L1 * l[2]);
      constexpr int sa = BV::sa, sb = BV::sb, sc = BV::sc;
      // face-point data, loaded before the staged batches are used
      R mf[9], vis = 0, cond = 0, dif[NS > 0 ? NS : 1];
      const unsigned pD_b = (NS > 0 && DIR == 2 && (S.quirks & CFX_Q_HV_RHOD_J))
                                ? (unsigned)(pl_i + B.sy) * (unsigned)sizeof(R) : pr_b;
      if (face_ok) {
#pragma unroll
        for (int rc = 0; rc < 9; ++rc) {
          if (rc / 3 == A) {  // the row of the face direction, kept from the face-area vector
            mf[rc] = mfa[rc % 3];
            continue;
          }
          if constexpr (DIR == 0) {
            R a, b;
            ld2(MT_ + rc * np, a, b);
            mf[rc] = (q5 && rc != 0) ? R(0.5) * (a + a) : R(0.5) * (a + b);
          } else {
            // the quirk's average (a + a) / 2 reads the left point twice: a
            // select of the address, not of the value, so that both loads
            // stay unconditional (a value select turns into a branch with
            // a wait for the first load)
            const R a = ldg<R>(MT_ + rc * np, pl_b);
            const R b = ldg<R>(MT_ + rc * np, (q5 && rc != 0) ? pl_b : pr_b);
            mf[rc] = R(0.5) * (a + b);
          }
        }
        R a, b;
        ld2(PR_ + 2 * np, a, b);
        vis = R(0.5) * (a + b);
        ld2(PR_ + 3 * np, a, b);
        cond = R(0.5) * (a + b);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) {
          if constexpr (DIR == 0) {  // the rho*D quirk reads j+1 only on z faces
            ld2(PR_ + (4 + s2) * np, a, b);
          } else {
            a = ldg<R>(PR_ + (4 + s2) * np, pl_b), b = ldg<R>(PR_ + (4 + s2) * np, pD_b);
          }
          dif[s2] = R(0.5) * (a + b);
        }
      } else {
#pragma unroll
        for (int rc = 0; rc < 9; ++rc) mf[rc] = R(0);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) dif[s2] = R(0);
      }
      auto chain = [&](const R gc[3], R out[3]) {
        out[0] = gc[0] * mf[0] + gc[1] * mf[3] + gc[2] * mf[6];
        out[1] = gc[0] * mf[1] + gc[1] * mf[4] + gc[2] * mf[7];
        out[2] = gc[0] * mf[2] + gc[1] * mf[5] + gc[2] * mf[8];
      };
      R dvel[4][3], avg[4];
      R dp[NS > 0 ? NS : 1], yb[NS > 0 ? NS : 1];
      R ucx = 0, ucy = 0, ucz = 0, e4 = 0;
      int vrg = kRangeMixed;  // NASA-7 range of the face temperatures of the wavefront
      static_for<NVB>([&](auto bi) {
        constexpr int nb2 = decltype(bi)::value;
        constexpr int n = NI + nb2;
        constexpr int j0 = nb2 * KV;
        constexpr int K = NVV - j0 < KV ? NVV - j0 : KV;
        R *const cur = n % 2 == 0 ? buf0 : buf1;
        R *const nxt = n % 2 == 0 ? buf1 : buf0;
        land(none{});
        if constexpr (nb2 + 1 < NVB) issue_vis(std::integral_constant<int, nb2 + 1>{}, nxt);
        static_for<K>([&](auto kk) {
          constexpr int k = decltype(kk)::value;
          constexpr int j = j0 + k;
          const R *q = cur + k * VPV + li;
          R gc[3], d[3];
          gc[A] = q[sa] - q[0];
          gc[Bx] = R(0.25) * (q[sb] - q[-sb] + q[sa + sb] - q[sa - sb]);
          gc[Cx] = R(0.25) * (q[sc] - q[-sc] + q[sa + sc] - q[sa - sc]);
          chain(gc, d);
          if constexpr (j < 4) {
            dvel[j][0] = d[0], dvel[j][1] = d[1], dvel[j][2] = d[2];
            avg[j] = R(0.5) * (q[0] + q[sa]);
            if constexpr (j == 3) {
              const R *du = dvel[0], *dv = dvel[1], *dw = dvel[2], *dT = dvel[3];
              const R third = R(1.0 / 3.0);
              const R txx = vis * (R(4) * du[0] - R(2) * dv[1] - R(2) * dw[2]) * third;
              const R tyy = vis * (R(4) * dv[1] - R(2) * du[0] - R(2) * dw[2]) * third;
              const R tzz = vis * (R(4) * dw[2] - R(2) * du[0] - R(2) * dv[1]) * third;
              const R txy = vis * (du[1] + dv[0]);
              const R txz = vis * (du[2] + dw[0]);
              const R tyz = vis * (dv[2] + dw[1]);
              const R G1 = k1 * txx + k2 * txy + k3 * txz;
              const R G2 = k1 * txy + k2 * tyy + k3 * tyz;
              const R G3 = k1 * txz + k2 * tyz + k3 * tzz;
              const R G4 = avg[0] * G1 + avg[1] * G2 + avg[2] * G3 + cond * (k1 * dT[0] + k2 * dT[1] + k3 * dT[2]);
              fbuf[1 * NT + tid] -= G1;
              fbuf[2 * NT + tid] -= G2;
              fbuf[3 * NT + tid] -= G3;
              fbuf[4 * NT + tid] -= G4;
              if constexpr (NS > 0) vrg = thermo_range<R>(sp, avg[3], avg[3]);
            }
          } else {
            constexpr int s2 = j - 4;
            yb[s2] = R(0.5) * (q[0] + q[sa]);
            R h;
            if (vrg == kRangeHi) h = species_h<R, kRangeHi>(sp, s2, avg[3]);
            else if (vrg == kRangeLo) h = species_h<R, kRangeLo>(sp, s2, avg[3]);
            else h = species_h<R>(sp, s2, avg[3]);
            const R proj = k1 * d[0] + k2 * d[1] + k3 * d[2];
            e4 += dif[s2] * h * proj;
            ucx += dif[s2] * d[0];
            ucy += dif[s2] * d[1];
            ucz += dif[s2] * d[2];
            dp[s2] = dif[s2] * proj;
          }
        });
      });
      if constexpr (NS > 0) {
        const R uc = k1 * ucx + k2 * ucy + k3 * ucz;
        fbuf[4 * NT + tid] -= e4;
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) fbuf[(5 + s2) * NT + tid] -= dp[s2] - yb[s2] * uc;
      }
    }

    // ---- cell difference of the face slots, dq written once per direction;
    // the dq being accumulated into is read before the barrier, so that its
    // latency overlaps the wait for the other wavefronts
    int c[3] = {g[0], g[1], g[2]};
    c[A] += 1;
    const unsigned qb = cell_ok ? (unsigned)((c[2] * B.my + c[1]) * B.mx + c[0]) * (unsigned)sizeof(R) : 0u;
    const gptr<R> __restrict__ DQ = Fd::dq(B);
    R dq0[NV];
    if (accumulate && cell_ok) {
#pragma unroll
      for (int v = 0; v < NV; ++v) dq0[v] = ldg<R>(DQ + v * B.ni, qb);
    }
    __syncthreads();
    if (cell_ok) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const R d = -(fbuf[v * NT + tid + ta_stride] - fbuf[v * NT + tid]);
        const gptr<R> dst = atg<R>(DQ + v * B.ni, qb);
        *dst = accumulate ? dq0[v] + d : d;
      }
    }
  }
}

// One workgroup per tile (PERSIST = false), or a grid smaller than the tile
// count whose workgroups stay resident and stride through the sequence
// (PERSIST = true; a separate kernel because the loop costs registers).
template <int NS, int DIR, int TA, int TB, int TC, int WPE, typename R, int G, bool PERSIST>
__global__ void __launch_bounds__(TA *TB *TC) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    rhs_pencil_kernel(const BlockView *__restrict__ views, const int *__restrict__ off, int nb, int padded,
                      const SpeciesDev *__restrict__ spg, SchemeDev S, int accumulate) {
  using P = PencilShape<NS, DIR, TA, TB, TC, R, WPE>;
  __shared__ R stg0[P::STG], stg1[P::STG];  // two batches of staged variables
  __shared__ R fbuf[P::NV * P::NT];    // each thread's face flux F_inv - F_visc
  const SpeciesPtrT<R> sp = const_table<R>(spg);
  if constexpr (PERSIST) {
    for (int t = blockIdx.x; t < padded; t += gridDim.x)
      rhs_pencil_tile<NS, DIR, TA, TB, TC, R, G, WPE>(views, off, nb, padded, t, sp, S, accumulate, stg0, stg1, fbuf);
  } else {
    rhs_pencil_tile<NS, DIR, TA, TB, TC, R, G, WPE>(views, off, nb, padded, blockIdx.x, sp, S, accumulate, stg0, stg1, fbuf);
  }
}

}  // namespace cfx
