// Gridding phase 1 for 3D on the matrix cores (v_mfma_f32_16x16x4f32).
//
// The KB weight of sample s at patch point (p0, p1, p2) is separable:
// w0[p0] * w1[p1] * w2[p2]. The plan orders a tile's samples by "cell": corner plane l0,
// pair of rows l1 / 2 and group of four positions l2 / 4 (bin.hip, cell_major_local), so
// the samples of a cell are contiguous and all land in one region of
// P x (P + 1) x (P + 3) subgrid points. Their sum over the region is a matrix product
// with K = samples:
//
//   R[(p2, coil, re/im), (p0, p1)] += sum_s  (w2_s[p2] * y_s[coil].re/im) * (w0_s[p0] w1_s[p1])
//
// with w1, w2 placed at the sample's offset inside the cell. 16 rows = 8/NCW values of
// p2 x NCW coils x (re, im); 16 columns = 16 (p0, p1) pairs; one instruction takes 4
// samples. A wave owns NCW coils and keeps R in registers while it walks a cell. The
// ring values of the cell's region are read before the cell's last group of products is
// issued and written back (old + R) right after it, so the read latency hides behind the
// matrix work; each lane's 4 result values (re and im of 2 coils at one p2: 16 contiguous
// bytes) take one 128-bit read and one write per 16x16 block. Compared
// with one LDS read-modify-write of P^3 points per sample and coil, LDS traffic drops by
// the samples per cell (18 on average for a 3D kooshball) and the multiply-adds move to
// the matrix cores.
//
// Plane ring and scratch format are those of grid_partial_kernel: a cell at plane l0
// touches planes l0 .. l0+P-1; when the plane advances the planes below it are stored
// (dense, per coil SUB^3 points) to the item's scratch slot and their ring slots cleared.
//
// Block: NW waves (NCW*NW coils) per (item, coil group). The waves share an LDS stage of
// CH samples (B operands, dim-2 weights) and each stages its coils' k-space values.
#pragma once

#ifndef MRA_MFMA_NCW
#define MRA_MFMA_NCW 2  // coils per wave (2 or 4)
#endif
#ifndef MRA_MFMA_NW
#define MRA_MFMA_NW 2   // waves per block (1, 2 or 4)
#endif
#ifndef MRA_MFMA_E1
#define MRA_MFMA_E1 2   // cell extent along dim 1 (corner rows: 2, 4 or 8)
#endif
#ifndef MRA_MFMA_E2
#define MRA_MFMA_E2 4   // cell extent along dim 2 (corner positions: 4 or 8)
#endif
#ifndef MRA_MFMA_KS
#define MRA_MFMA_KS 4   // matrix instructions' K-steps (4 samples each) per control step
#endif

namespace mra {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int P, int NWV = MRA_MFMA_NW, int E1V = MRA_MFMA_E1, int PEV = P>
struct MfmaGrid {
  static constexpr int NP = P;                 // patch points per dim of the plan's weights
  static constexpr int PE = PEV;               // patch points per dim the samples reach
  static constexpr int T = 8;                  // 3D tile edge
  static constexpr int SUB = T + P - 1;        // subgrid edge
  static constexpr int PP = SUB * SUB;         // dense points per plane
  static constexpr int SUBN = SUB * PP;
  static constexpr int NCW = MRA_MFMA_NCW;     // coils per wave
  static constexpr int NW = NWV;               // waves per block
  static constexpr int NC = NW * NCW;          // coils per block
  static constexpr int PV = 8 / NCW;           // values of p2 per 16 rows
  static constexpr int PF = 2 * NCW;           // floats per grid point in the ring
  static constexpr int E1 = E1V;               // cell extent along dim 1 (corner rows)
  static constexpr int E2 = MRA_MFMA_E2;       // cell extent along dim 2 (corner positions)
  static constexpr int E1P = E1 + PE - 1;      // region extent along dim 1
  static constexpr int E2P = E2 + PE - 1;      // region extent along dim 2
  static constexpr int RB = (E2P + PV - 1) / PV;  // row blocks
  static constexpr int CWS = RB;               // floats per (sample, p2 mod PV) of dim-2 weights
  static constexpr int NJ = PE * E1P;          // (p0, p1) columns
  static constexpr int CB = (NJ + 15) / 16;    // column blocks
  static constexpr int CPW = CB * 16 / NW;     // B columns staged per wave
  static constexpr int CH = 64;                // samples per stage
  static constexpr int KS = MRA_MFMA_KS;       // K-steps per control step (up to 4*KS samples of a cell)
  // ring layout (floats): [slot P][p1 SUB][p2 SUB][coil NCW][re, im]; row and plane
  // pitches padded (a bank model of the cell flush: at most 2-way conflicts)
  static constexpr int ROW = SUB * PF + 4;
  static constexpr int PLANE = SUB * ROW + (NCW == 2 ? (E1 == 4 ? 48 : 8) : 40);
  static constexpr int DUM = 64 * 4 + RB * PV * PF;  // dummy slots of lanes with no result
  static constexpr int RING = P * PLANE + DUM;
  static_assert(NCW == 2 || NCW == 4, "2 or 4 coils per wave");
  static_assert(NW == 1 || NW == 2 || NW == 4, "1, 2 or 4 waves per block");
  static_assert(CPW % 4 == 0, "B columns are staged four at a time");
  static_assert(PE >= 1 && PE <= P, "patch points reached");
  // cells must be contiguous in the plan's order (a0, a1 / 2, a2 / 4, a1 % 2, a2 % 4)
  static_assert((E1 == 1 && E2 == 4) || (E1 == 2 && (E2 == 4 || E2 == 8)) || (E2 == 8 && (E1 == 4 || E1 == 8)),
                "cell shape");
};

// Planes lo .. hi-1 of this wave's ring are final: store them to scratch (dense) and
// clear their slots. Planes the ring never held (p >= lo + P) are stored as zeros.
template <class M>
__device__ __forceinline__ void mfma_retire(int lo, int hi, int nc, float* ring, float2* __restrict__ dst) {
  constexpr int P = M::NP;
  constexpr int NE = M::NCW * M::PP;
  constexpr int NIT = (NE + 63) / 64;
  const int lane = threadIdx.x & 63;
  for (int p = lo; p < hi; ++p) {
    const bool held = p < lo + P;
    float* slot = ring + (p % P) * M::PLANE;
    float2 v[NIT];
    float2* src[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = lane + 64 * it;
      const int c = e / M::PP, r = e - c * M::PP;
      const int p1 = r / M::SUB, p2 = r - p1 * M::SUB;
      src[it] = (float2*)(slot + p1 * M::ROW + p2 * M::PF + 2 * c);
      v[it] = (held && e < NE) ? *src[it] : make_float2(0.f, 0.f);
    }
    if (held) {
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        if (lane + 64 * it < NE) *src[it] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = lane + 64 * it;
      const int c = e / M::PP, r = e - c * M::PP;
      if (e < NE && c < nc) dst[(long long)c * M::SUBN + (long long)p * M::PP + r] = v[it];
    }
  }
}

// B operand columns W*CPW .. (W+1)*CPW-1 of this lane's sample: column j = (p0, p1) =
// (j / E1P, j % E1P) holds w0[p0] * w1[p1 - off1] (0 outside the patch or past NJ).
// Layout [cb][sample][16]; four columns per 128-bit store.
template <class M, int W>
__device__ __forceinline__ void mfma_stage_b(const float (&rw)[3 * M::NP], int off1, float* sbw, int lane) {
  constexpr int P = M::NP;
#pragma unroll
  for (int c4 = 0; c4 < M::CPW / 4; ++c4) {
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = W * M::CPW + 4 * c4 + u;
      const int p0 = j / M::E1P, p1 = j % M::E1P;
      float w1 = 0.0f;
#pragma unroll
      for (int o = 0; o < M::E1; ++o)
        if (p1 - o >= 0 && p1 - o < M::PE) w1 = off1 == o ? rw[P + p1 - o] : w1;
      v[u] = j < M::NJ ? rw[p0] * w1 : 0.0f;
    }
    const int j = W * M::CPW + 4 * c4;
    *(v4f*)(sbw + ((j >> 4) * M::CH + lane) * 16 + (j & 15)) = v4f{v[0], v[1], v[2], v[3]};
  }
}

// Ring addresses of this lane's result values for one cell: rows 4q..4q+3 of block
// (rb, cb) are p2 = c2 + PV rb + 4q / PF, floats (4q mod PF) .. +3 of that point; column
// (p0, p1) = (j / E1P, j % E1P), j = 16 cb + lane % 16. Columns past NJ and rows past E2P
// carry zeros and go to the lane's dummy slot.
template <class M>
struct CellFlush {
  static constexpr int P = M::NP;
  v4f* ptr[M::RB][M::CB];
  v4f old[M::RB][M::CB];

  // lane constants: j0 (-1 past NJ), offset of (p1 - l1, p2 - c2, float) in the ring
  __device__ __forceinline__ void locate(float* ring, int ckey, const int (&j0)[M::CB], const int (&loff)[M::CB],
                                         int prow, int lane) {
    const int pslot = (ckey >> 16) % P;
    const int cbase = ((ckey >> 8) & 0xff) * M::E1 * M::ROW + (ckey & 0xff) * M::E2 * M::PF;
    float* dummy = ring + P * M::PLANE + 4 * lane;
#pragma unroll
    for (int cb = 0; cb < M::CB; ++cb) {
      int slot = pslot + j0[cb];
      slot = slot >= P ? slot - P : slot;
      float* base = j0[cb] >= 0 ? ring + slot * M::PLANE + loff[cb] + cbase : dummy;
#pragma unroll
      for (int rb = 0; rb < M::RB; ++rb) {
        float* pt = base + rb * M::PV * M::PF;
        if constexpr (M::E2P % M::PV != 0) pt = M::PV * rb + prow < M::E2P ? pt : dummy + rb * M::PV * M::PF;
        ptr[rb][cb] = (v4f*)pt;
      }
    }
  }
  __device__ __forceinline__ void read() {
#pragma unroll
    for (int cb = 0; cb < M::CB; ++cb)
#pragma unroll
      for (int rb = 0; rb < M::RB; ++rb) old[rb][cb] = *ptr[rb][cb];
  }
  __device__ __forceinline__ void write(const v4f (&acc)[M::RB][M::CB]) {
#pragma unroll
    for (int cb = 0; cb < M::CB; ++cb)
#pragma unroll
      for (int rb = 0; rb < M::RB; ++rb) *ptr[rb][cb] = old[rb][cb] + acc[rb][cb];
  }
};

// One (item, coil group) with configuration M: the work of grid_mfma_kernel below. The LDS
// buffers come from the kernel, sized for every configuration it may pick.
template <class M>
__device__ __forceinline__ void grid_mfma_item(const float2* __restrict__ ksp, const float* __restrict__ dcf,
                                               const float* __restrict__ wsort, const int* __restrict__ lbase,
                                               const int* __restrict__ perm, const Geom& g, const Item& it,
                                               int iidx, int grp, float2* __restrict__ scratch, float* ring_all,
                                               float* sbw, float* scw, float* sy) {
  constexpr int P = M::NP;
  constexpr int SUB = M::SUB, CH = M::CH, RB = M::RB, CB = M::CB, PF = M::PF;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cw0 = grp * M::NC + wave * M::NCW;  // this wave's first coil
  const int cnt = it.w;
  const long long gs0 = g.samp_off + it.z;
  float* syw = sy + wave * CH * PF;
  float* ring = ring_all + wave * M::RING;
  float2* dst = scratch + g.scratch_off + ((long long)(iidx - g.item_base) * g.nbatch + cw0) * M::SUBN;
  const int nc = max(0, min(M::NCW, g.nbatch - cw0));  // this wave's valid coils
  const int q = lane >> 4;         // sample slot k of the operands; row group of the result
  const int i16 = lane & 15;       // row of A / column of B and of the result
  const int p2s = i16 / PF;        // A row: p2 mod PV
  const int yi = i16 % PF;         // A row: coil * 2 + re/im
  const int prow = (4 * q) / PF;   // result rows: p2 offset and float offset in the point
  int j0[CB], loff[CB];
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    const int j = cb * 16 + i16;
    j0[cb] = j < M::NJ ? j / M::E1P : -1;
    loff[cb] = (j % M::E1P) * M::ROW + prow * PF + (4 * q) % PF;
  }

  // gathered sample data of the next stage (registers); the sample indices (perm) are
  // read one stage earlier still, so no gather waits for a dependent load
  int rbase = 0;
  float rw[3 * P];
  float2 ry[M::NCW];
  int onext = lane < cnt ? sample_slot(perm, g, gs0 + lane) : 0;
  auto gather = [&](int s0) {
    const bool ok = s0 + lane < cnt;
    const long long gs = gs0 + (ok ? s0 + lane : 0);
    const int o = onext;
    rbase = ok ? lbase[gs] : 0;
    const float* wp = wsort + gs * (3 * P);
#pragma unroll
    for (int e = 0; e < 3 * P; ++e) rw[e] = ok ? wp[e] : 0.0f;
    const float sc = (ok && dcf) ? dcf[g.dcf_plan_order ? gs : g.samp_off + o] * g.out_scale : g.out_scale;
    const float2* kc = ksp + g.ksp_off + (long long)cw0 * g.ksp_stride + o;
#pragma unroll
    for (int c = 0; c < M::NCW; ++c) {
      const float2 v = (ok && c < nc) ? kc[(long long)c * g.ksp_stride] : make_float2(0.f, 0.f);
      ry[c] = make_float2(v.x * sc, v.y * sc);
    }
    onext = s0 + CH + lane < cnt ? sample_slot(perm, g, gs0 + s0 + CH + lane) : 0;
  };
  gather(0);

  for (int e = lane; e < M::RING / 4; e += 64) ((v4f*)ring)[e] = v4f{0.f, 0.f, 0.f, 0.f};

  v4f acc[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
  int lo = 0;            // lowest plane not yet stored (the ring holds lo .. lo+P-1)
  int ckey = 0;          // active cell: l0 << 16 | (l1 / E1) << 8 | l2 / E2
  bool pending = false;  // acc holds a cell not yet written back (it ran past a stage end)
  int prevkey = -1;      // cell of the last sample of the previous stage

  for (int s0 = 0; s0 < cnt; s0 += CH) {
    const int n = min(CH, cnt - s0);
    // stage this chunk: every wave writes its share of the shared operands
    int key;
    {
      const int l2 = rbase % SUB, l12 = rbase / SUB;
      const int l1 = l12 % SUB, l0 = l12 / SUB;
      key = (l0 << 16) | ((l1 / M::E1) << 8) | (l2 / M::E2);
      const int off1 = l1 % M::E1, off2 = l2 % M::E2;
      switch (wave) {
        case 0: mfma_stage_b<M, 0>(rw, off1, sbw, lane); break;
        case 1: mfma_stage_b<M, 1 % M::NW>(rw, off1, sbw, lane); break;
        case 2: mfma_stage_b<M, 2 % M::NW>(rw, off1, sbw, lane); break;
        default: mfma_stage_b<M, 3 % M::NW>(rw, off1, sbw, lane); break;
      }
      if (wave == M::NW - 1) {
        // dim-2 weights on the region: cw[p2] = w2[p2 - off2], stored by p2 mod PV
#pragma unroll
        for (int ps = 0; ps < M::PV; ++ps)
#pragma unroll
          for (int rb = 0; rb < M::CWS; ++rb) {
            const int t = M::PV * rb + ps - off2;
            float v = 0.0f;
#pragma unroll
            for (int u = 0; u < M::PE; ++u) v = t == u ? rw[2 * P + u] : v;
            scw[(lane * M::PV + ps) * M::CWS + rb] = v;
          }
      }
#pragma unroll
      for (int c = 0; c < M::NCW; ++c) {
        syw[lane * PF + 2 * c] = ry[c].x;
        syw[lane * PF + 2 * c + 1] = ry[c].y;
      }
    }
    __syncthreads();
    if (s0 + CH < cnt) gather(s0 + CH);  // lands while this chunk is accumulated

    // cell starts (smask) and group starts (gmask: every 4th sample from the last cell
    // start, or from the stage start for a cell carried over) of this chunk, uniform
    const int kprev = __shfl_up(key, 1);
    const bool st = lane < n && (lane == 0 ? key != prevkey : key != kprev);
    const unsigned long long smask = __ballot(st);
    prevkey = __builtin_amdgcn_readlane(key, n - 1);
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
    const unsigned long long sb = smask & upto;
    const int cs = sb ? 63 - __builtin_clzll(sb) : 0;
    unsigned long long gm = __ballot(lane < n && ((lane - cs) % (4 * M::KS)) == 0);

    // raw operands of a group of gsz <= 4*KS samples at s: per K-step k (samples
    // s+4k .. s+4k+3) y, dim-2 weights, B columns
    constexpr int KS = M::KS;
    struct Ops {
      float y[KS], cw[KS][RB], bv[KS][CB];
    };
    auto load_ops = [&](Ops& o, int s, int gsz) {
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        if (4 * k < gsz) {
          const int sk = s + 4 * k + (4 * k + q < gsz ? q : 0);  // empty slot: zero A, B of a valid sample
          o.y[k] = 4 * k + q < gsz ? syw[sk * PF + yi] : 0.0f;
          const float* cwp = scw + (sk * M::PV + p2s) * M::CWS;
#pragma unroll
          for (int rb = 0; rb < RB; ++rb) o.cw[k][rb] = cwp[rb];
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) o.bv[k][cb] = sbw[(cb * CH + sk) * 16 + i16];
        }
      }
    };
    int s = (int)__builtin_ctzll(gm);  // 0
    gm &= gm - 1;
    int s2 = gm ? (int)__builtin_ctzll(gm) : n;
    // One group with operands `cur`; the next group's operands go to `nxt` once this
    // group's products are issued. Returns false after the chunk's last group. The loop
    // below alternates two operand sets so that no register copy waits for the loads.
    auto step = [&](Ops& cur, Ops& nxt) -> bool {
      const unsigned long long gm2 = gm & (gm - 1);
      const int s3 = gm2 ? (int)__builtin_ctzll(gm2) : n;
      if ((smask >> s) & 1ull) {
        // a new cell; the previous one was written back unless it ran to a stage end
        const int nk = __builtin_amdgcn_readlane(key, s);
        if (pending) {
          CellFlush<M> f;
          f.locate(ring, ckey, j0, loff, prow, lane);
          f.read();
          f.write(acc);
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};
        const int pl = nk >> 16;
        if (pl > lo) {
          mfma_retire<M>(lo, pl, nc, ring, dst);
          lo = pl;
        }
        ckey = nk;
        pending = true;
      }
      const bool last = s2 < n && ((smask >> s2) & 1ull);  // the cell ends with this group
      // written out for both paths so that each waits only for its own loads
      const int gsz = s2 - s;
      auto products = [&]() {
#pragma unroll
        for (int k = 0; k < KS; ++k) {
          if (4 * k < gsz) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
              const float av = cur.cw[k][rb] * cur.y[k];
#pragma unroll
              for (int cb = 0; cb < CB; ++cb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, cur.bv[k][cb], acc[rb][cb], 0, 0, 0);
            }
          }
        }
        if (s2 < n) load_ops(nxt, s2, s3 - s2);
      };
      if (last) {
        CellFlush<M> f;
        f.locate(ring, ckey, j0, loff, prow, lane);
        f.read();
        products();
        f.write(acc);
        pending = false;
      } else {
        products();
      }
      if (s2 >= n) return false;
      s = s2;
      s2 = s3;
      gm = gm2;
      return true;
    };
    Ops opa, opb;
    load_ops(opa, s, s2 - s);
    while (step(opa, opb) && step(opb, opa)) {
    }
    __syncthreads();  // the stage is rewritten next
  }
  if (pending) {
    CellFlush<M> f;
    f.locate(ring, ckey, j0, loff, prow, lane);
    f.read();
    f.write(acc);
  }
  mfma_retire<M>(lo, SUB, nc, ring, dst);
}


// Kernel: one block per (item, coil group). Items whose samples never reach the last patch
// point of a dimension (their weight there is exactly 0: t - w/2 off the grid integers, every
// sample but the k-space centre's for integer widths) run with P - 1 points per dimension and
// cells of 1 x 4 corners: at P = 5 the (p0, p1) columns are 4 x 4 = 16, one matrix block
// instead of two, so a sample costs half the matrix instructions and B operands. Items with
// such a sample keep P points and 2 x 4 cells. The decision is per item, from one pass over
// the last weights of its samples (L2-resident: the item's coil groups run back to back on
// one XCD). Both configurations write the same scratch subgrids.
#ifndef MRA_MFMA_SHORT
#define MRA_MFMA_SHORT 1
#endif
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int P, int NWV>
__global__ __launch_bounds__(64 * NWV) void grid_mfma_kernel(
    const float2* __restrict__ ksp, const float* __restrict__ dcf, const float* __restrict__ wsort,
    const int* __restrict__ lbase, const int* __restrict__ perm, const Geom* __restrict__ geoms,
    const Item* __restrict__ items, int nitems, int ngroups, float2* __restrict__ scratch) {
  using MF = MfmaGrid<P, NWV>;
  constexpr bool kShort = MRA_MFMA_SHORT && P >= 4;
  using MS = MfmaGrid<P, NWV, kShort ? 1 : MRA_MFMA_E1, kShort ? P - 1 : P>;
  static_assert(MF::NC == MS::NC, "same coils per block");
  __shared__ __attribute__((aligned(16))) float ring_all[cmax(MF::NW * MF::RING, MS::NW * MS::RING)];
  __shared__ __attribute__((aligned(16))) float sbw[cmax(MF::CB, MS::CB) * MF::CH * 16];  // B operands
  __shared__ __attribute__((aligned(16))) float scw[MF::CH * MF::PV * cmax(MF::CWS, MS::CWS)];  // dim-2 weights
  __shared__ __attribute__((aligned(16))) float sy[MF::NW * MF::CH * MF::PF];             // k-space

  int iidx, grp;
  item_of_block(blockIdx.x, ngroups, iidx, grp);
  if (iidx >= nitems) return;
  const Item it = items[iidx];
  if (it.w <= 0 || it.y < 0) return;  // non-finite samples contribute nothing
  const Geom g = geoms[it.x];
  if (grp * MF::NC >= g.nbatch) return;
  bool full = true;
  if constexpr (kShort) {
    const float* w0 = wsort + (g.samp_off + it.z) * (3 * P);
    int edge = 0;
    for (int s = threadIdx.x; s < it.w; s += 64 * NWV) {
      const float* w = w0 + (long long)s * (3 * P);
      edge |= (w[P - 1] != 0.0f) | (w[2 * P - 1] != 0.0f) | (w[3 * P - 1] != 0.0f);
    }
    full = __syncthreads_or(edge) != 0;
  }
  if (full)
    grid_mfma_item<MF>(ksp, dcf, wsort, lbase, perm, g, it, iidx, grp, scratch, ring_all, sbw, scw, sy);
  else
    grid_mfma_item<MS>(ksp, dcf, wsort, lbase, perm, g, it, iidx, grp, scratch, ring_all, sbw, scw, sy);
}

}  // namespace mra
