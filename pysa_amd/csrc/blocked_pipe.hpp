// The pipelined kernel of the MFMA form (a template over the coupling
// encoding, spin words per matrix wave, the visit order and the tiles per
// workgroup), instantiated one encoding per translation unit.
#pragma once
#include "blocked_dev.hpp"

namespace pysa {
namespace blk {

// ------------------------------------------------------- pipelined form --
// Eight waves per group of 32 NT replicas (NT tiles of 32), two per SIMD.
// Wave 0 runs the proposals with one lane per replica: the lane holds all 32
// fields of the block, so a decision needs no cross-lane broadcast.  Three
// matrix waves (SIMDs 1-3), a third of the K range each and every tile,
// compute the next block's fields over every K block except c (whose spins
// are changing) while wave 0 sweeps block c:
//     P_{c+1} = sum_{b != c} J[c+1, b] . S_b
// When wave 0 is done with block c it adds the missing product
// J[c+1, c] . S_c itself (one K block on SIMD 0's matrix core), so the fields
// of block c+1 are ready at the block's barrier.  The four helper waves stage
// the next block's J block, h and row sums and compute its acceptance
// brackets (those on SIMDs 1-3), so the chain only compares, flips and
// updates.  Shared state alternates between two buffers by the parity of a
// running block counter.
// K ranges of the pipelined kernel (one matrix wave each).
constexpr int kPipeMW = 3;
// What the pipelined kernel runs (its MODE).
constexpr int kModePT = 0, kModeBatched = 1, kModeAIS = 2;
constexpr int kPipePF = 3;  // fragment stages in flight per matrix wave (slot ring)

// A replica's 32 fields take 32 floats; its 16-byte row groups are XOR-
// swizzled by the replica index so that 8 lanes of a 16-byte access hit all
// 32 banks.
__device__ __forceinline__ void swap2(float& a, float& b) {
  const float t = a;
  a = b;
  b = t;
}
// Field layout [replica][row]: the 16-byte groups of four rows are XOR
// swizzled by the replica so that a 16-byte access by consecutive replicas
// hits distinct banks in every aligned run of 8 lanes (32 banks) and of 16
// lanes (64 banks: consecutive rows of 128 bytes alternate bank halves).
__device__ __forceinline__ int fld(int q, int r) {
  return q * 32 + 4 * ((r >> 2) ^ (q & 7) ^ ((q >> 3) & 1)) + (r & 3);
}
constexpr int kHL = 36;
constexpr int kFG = 2;  // row groups per stage of the proposal wave's field sums
// NI: instances per workgroup (the batched form keeps one J block and h per
// tile).
template <int MW, int NI>
struct __align__(16) PipeLds {
  float part[2][MW][64 * 32];  // [parity][matrix wave][fld(replica, row)]
  float corr[64 * 32];         // J[c+1, c] . S_c of the next block
  float jb[2][NI][kB * kB];   // row i of J[blk][blk] right of the diagonal, shifted left:
                              // J^T[pos0+i][pos0+i+1+t] at i*32 + t (t < 31 - i), else 0
  float hb[2][NI][kB];
  float rsb[2][kB];
  // acceptance brackets [parity][replica][visit index], rows padded to
  // kHL floats so that 16-byte reads of consecutive replicas hit distinct banks
  float lo[2][64][kHL];
  float hi[2][64][kHL];
  uint32_t bb[2][64];         // block spins before the block (bit k = row k ^ the visit mask)
  uint32_t bnew[2][64];       // block spins after the block (by block parity)
  float sfp[64 * 33];         // exact rerun staging
  // visit schedule by pass parity: block at step t | its in-block XOR mask << 8
  uint16_t sch[2][64];
  float E[64];
  float rb[64];
  float bet[64];  // the ladder by slot (batched: tile t's instance's at 32 t)
  float Es[64];   // energy of the replica in each slot, before the exchanges
  int bso[64];    // its ascending order (log-omega)
  int bi[64];
  int flag[64];
  int src[64];
  double lu[64];
};

// BR: block-random visit order (schedule tables in LDS); otherwise blocks are
// visited in order with no in-block shuffle, and every schedule entry is
// plain arithmetic on the step.
// BAT: the batched form.  Each workgroup runs two tiles (its entry of
// p.bat), each up to 32 chains of one instance of a batch, the two instances
// with the same number of blocks; couplings are scale * S (S int8), the
// fields kept as the exact integer counts S s, and the field of a spin is the
// f32 value scale * c + h, rounded twice like the batched chain of the
// oracle; the energy changes are summed in f64 in visit order.
// MODE: kModePT (Metropolis sweeps with parallel tempering), kModeBatched
// (BAT), kModeAIS (AIS: annealed importance sampling with the uniform
// reference; no exchanges, no best tracking; every chain at the schedule's
// beta of the sweep, its energy tracked in f64 and its log-weight updated
// after every sweep, pysa/simulation.py:293-317).
template <int NP, bool I8, int NW, int MW, bool BR, int NT, int MODE = kModePT>
__global__ __launch_bounds__(128 * (MW + 1)) void blocked_pipe_kernel(BParams p0) {
  constexpr bool BAT = MODE == kModeBatched, AIS = MODE == kModeAIS;
  // a local copy (the register allocation of the block loops was found to be
  // sensitive to this; the copy keeps it as measured)
  BParams p = p0;
  static_assert(!BAT || (I8 && NP == 1 && NT == 2),
                "the batched form multiplies int8 couplings, two tiles per workgroup");
  constexpr int NI = BAT ? NT : 1;  // instances per workgroup (coupling sets)
  __shared__ PipeLds<MW, NI> L;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tid = threadIdx.x;
  const int read0 = blockIdx.x * p.rpw;
  const int R = p.R, n = p.n;
  const int nblk = BAT ? p.bat[blockIdx.x].nblk : p.nblk;
  const int kb = BAT ? p.bat[blockIdx.x].kb : p.kb;
  // batched: tile tt's instance and reads, read once at the start (before
  // any store, so they are scalar loads kept in registers; a load later in
  // the kernel could not be proven unclobbered and would be a vector load)
  BatTile bt[NI];
  if constexpr (BAT) {
#pragma unroll
    for (int tt = 0; tt < NI; ++tt) bt[tt] = p.bat[blockIdx.x].t[tt];
  }
  auto tile = [&](int tt) -> const BatTile& { return bt[tt]; };
  auto n_of = [&](int tt) { return BAT ? tile(tt).n : n; };
  const int h = lane >> 5;
  // Two waves per SIMD (SIMD w & 3).  Wave 0 runs the proposals; matrix
  // wave w = 1..MW computes the products of K range m = w - 1 for both tiles
  // of 32 replicas (each coupling fragment is loaded once for both); the
  // helper waves MW+1.. stage the next blocks' J block, h and row sums,
  // compute the acceptance brackets (those on SIMDs 1..MW) and, wave MW+1,
  // the visit schedule.
  static_assert(MW == 3, "eight waves: the proposal wave, three matrix waves, four helpers");
  const bool mw = w >= 1 && w <= MW;  // a matrix wave
  const bool hw = w > MW;             // a helper wave
  const int m = mw ? w - 1 : 0;
  const int hl = hw ? (w - MW - 1) * 64 + lane : -1;  // helper-lane index 0..64 (MW+1) - 1
  // Operand role (matrix waves): replica qtt[tt] of tile tt.
  static_assert(NT == 1 || NT == 2, "one or two tiles of 32 replicas");
  // gtt: its read within the workgroup (recomputed where needed: kept in
  // registers across the block loop they would add to its pressure)
  int qtt[NT], gtt[NT];
  bool actt[NT];
  size_t base_t[NT];
  auto rtt_of = [&](int tt) {  // its replica within the read
    return BAT ? (lane & 31) % R : qtt[tt] - gtt[tt] * R;
  };
  auto rdt_of = [&](int tt) {  // the read's global index
    return BAT ? tile(tt).read0 + (lane & 31) / R : read0 + gtt[tt];
  };
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) {
    qtt[tt] = tt * 32 + (lane & 31);
    if constexpr (BAT) {  // tile tt: p.rpw reads of its own instance
      const int gl = (lane & 31) / R;
      gtt[tt] = tt * p.rpw + gl;
      actt[tt] = mw && gl < p.rpw && rdt_of(tt) < tile(tt).read_end;
      base_t[tt] = actt[tt] ? (size_t)(tile(tt).spin_adj +
                                       ((int64_t)rdt_of(tt) * R + rtt_of(tt)) * tile(tt).n)
                            : 0;
    } else {  // p.rpw reads over the tiles
      gtt[tt] = qtt[tt] / R;
      actt[tt] = mw && (gtt[tt] < p.rpw) && (read0 + gtt[tt] < p.reads);
      base_t[tt] = actt[tt] ? ((size_t)(read0 + gtt[tt]) * R + (qtt[tt] - gtt[tt] * R)) * n : 0;
    }
  }
  // the best state of tile tt's read (operand role)
  auto best_of = [&](int tt) -> int8_t* {
    if constexpr (BAT) return p.best_state + (tile(tt).best_adj + (int64_t)rdt_of(tt) * tile(tt).n);
    return p.best_state + (size_t)(read0 + gtt[tt]) * n;
  };
  // Replica role (every wave): lane = replica q of the group; g its read
  // within the workgroup, tq its tile in the batched form (whose instance's
  // size, scale and ladder the lane uses).
  const int q = lane;
  const int tq = BAT ? (q >> 5) : 0;
  const int gl = BAT ? (q & 31) / R : 0;  // batched: its read within the tile
  const int g = BAT ? tq * p.rpw + gl : q / R;
  const int rr = BAT ? (q & 31) - gl * R : q - g * R;
  const int read = BAT ? (tq ? tile(1).read0 : tile(0).read0) + gl : read0 + g;
  const bool act = BAT ? gl < p.rpw && read < (tq ? tile(1).read_end : tile(0).read_end)
                       : (g < p.rpw) && (read < p.reads);
  const int nq = BAT ? (tq ? tile(1).n : tile(0).n) : n;
  const float* betas_q =
      BAT ? p.betas + (size_t)(tq ? tile(1).inst : tile(0).inst) * R : p.betas;
  const uint64_t cid = (uint64_t)read * R + rr;
  const int c0 = mw ? m * kb : 0;
  const int jn = mw ? max(0, min(kb, nblk - c0)) : 0;

  Bits<NW> sa[NT];  // this wave's K range of each tile's spins, operand layout
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) {
    sa[tt].w = 0;
    for (int j = 0; j < jn; ++j) {
      const int c = c0 + j;
      uint32_t va = 0;
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const int rel = I8 ? (16 * h + b) : ((b < 8) ? (8 * h + b) : (16 + 8 * h + (b - 8)));
        const int pos = 32 * c + rel;
        if (pos < n_of(tt) && actt[tt] && p.spins[base_t[tt] + pos] == 1) va |= 1u << b;
      }
      set_half<NW>(sa[tt], j, va);
    }
  }
  float E = 0.f, bestE = 0.f, gE = 0.f;
  int gsw = 0, gimp = 0;
  unsigned long long acc = 0;
  double Ed = 0.0, lw = 0.0, bold = 0.0;  // AIS: f64 energy, log-weight, previous beta
  if (AIS && w == 0) {
    Ed = act ? p.ais_e[read] : 0.0;  // R = 1: chain = read
    L.bi[q] = 0;
    const float t0 = p.temps[0];
    L.rb[q] = act ? t0 : 0.f;
    // ordering 1: the weight of sweep 0 before it (bold = 0)
    if (p.ordering == 1 && act) lw = lw + ((-(double)t0 * Ed) - (-bold * Ed));
  }
  if (!AIS && w == 0) {
    E = act ? p.energies[(size_t)read * R + rr] : 0.f;
    L.bi[q] = act ? rr : 0;  // beta_idx = arange(R) (sa.py:325)
    L.rb[q] = act ? betas_q[rr] : 0.f;
    // the ladder in LDS for the exchanges (read serially by the read leaders)
    const int kq = BAT ? (q & 31) : q;
    L.bet[q] = kq < R ? betas_q[kq] : 0.f;
    L.bso[q] = (kq < R && p.logw && p.bsort) ? p.bsort[kq] : 0;
    bestE = E;
    gE = E;  // simulation.py:205 (meaningful on read leaders, rr == 0)
    if (p.best_sem == 0 && act) {
      p.best_rep_e[(size_t)read * R + rr] = E;
      p.best_rep_sw[(size_t)read * R + rr] = 0;
    }
  }
  // Fields of replicas no tile covers (NT = 1: lanes 32-63 of the proposal
  // wave) stay zero, so their idle decisions see finite values.
  for (int i = tid; i < 64 * 32; i += 128 * (MW + 1)) L.corr[i] = 0.f;
  for (int i = tid; i < 2 * MW * 64 * 32; i += 128 * (MW + 1)) (&L.part[0][0][0])[i] = 0.f;

  constexpr int PF = kPipePF;
  const uint32_t emul = p.ptype == 0 ? 0xFEu : 1u;
  const uint32_t exr = p.ptype == 0 ? 0xFFFFFFFFu : 0u;
  // operand-expansion constants in vector registers (see spins_bf16_r)
  uint32_t vmask = 0x80008000u, vone = 0x3F803F80u;
  asm volatile("" : "+v"(vmask), "+v"(vone));
  constexpr int NST = I8 ? 1 : 2;
  constexpr int NHL = 64 * (MW + 1);  // helper lanes
  constexpr int NJB = (kB * kB + NHL - 1) / NHL;
  typedef typename std::conditional<I8, i32x4, bf16x8>::type Frag;
  Frag ring[NI][PF][NST][NP];  // by instance (tile in the batched form)
  float jb_reg[NI][NJB], h_reg[NI] = {}, rs_reg = 0.f;
  uint32_t jb_ok[NI] = {};  // bit e2: jb_reg[.][e2] is a coupling; bit 31: h_reg / rs_reg are
  // the coupling fragments of instance ti
  auto frag_base = [&](int ti) -> const Frag* {
    if constexpr (BAT)
      return reinterpret_cast<const Frag*>(reinterpret_cast<const int8_t*>(p.frag) + tile(ti).frag);
    return reinterpret_cast<const Frag*>(p.frag);
  };
  auto frag_ptr = [&](int ti, int c, int j, int st, int pc) -> const Frag* {
    if (I8) return frag_base(ti) + ((size_t)c * nblk + c0 + j) * 64 + lane;
    return frag_base(ti) + (((size_t)c * (2 * nblk) + 2 * (c0 + j) + st) * NP + pc) * 64 + lane;
  };
  // Stage j of this wave's K range; stages past the range reload the
  // range's last K block (recently read, so a cache hit), so every path
  // issues the same loads and the wait counts stay exact (a skipped load
  // would make the compiler wait for all of them).
  // One K block of fragments: NST * NP * 64 of them, consecutive.
  constexpr int kStage = NST * NP * 64;
  auto load_stage = [&](int c, int j, Frag (&dst)[NST][NP], int ti) {
    // jn passes through an empty asm so that the per-stage offsets are
    // recomputed (a few scalar ops) rather than hoisted out of the block
    // loop into registers that would spill
    int jnl = __builtin_amdgcn_readfirstlane(jn);
    asm volatile("" : "+s"(jnl));
    const int jc = jnl > 0 ? min(j, jnl - 1) : nblk - 1 - c0;
    // wave-uniform base of the stage, then a per-lane 16-byte offset
    const Frag* base = frag_ptr(ti, c, 0, 0, 0) - lane + (size_t)jc * kStage;
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) dst[st][pc] = base[(st * NP + pc) * 64 + lane];
  };
  // Request row block c: the first PF fragment stages of this wave's K range
  // (matrix waves).
  auto prefetch_frags = [&](int c) {
#pragma unroll
    for (int u = 0; u < PF; ++u)
#pragma unroll
      for (int ti = 0; ti < NI; ++ti) load_stage(c, u, ring[ti][u], ti);
  };
  // Request row block c, visited with in-block mask xm, spread over the
  // helper lanes: J[blk][blk] in visit order, h and row sums.
  auto prefetch_jb = [&](int c, uint32_t xm) {
    const int pos0 = 32 * c;
    // Every lane loads (an in-bounds address when its element is padding)
    // and publish_block zeroes the padding: no register is written here
    // except by these loads, which keeps them from waiting on older loads.
    if constexpr (!BAT) {
      jb_ok[0] = 0;
#pragma unroll
      for (int e2 = 0; e2 < NJB; ++e2) {
        const int e = hl + e2 * NHL;
        // slot t = e & 31 of visit row i holds visit column k = i + 1 + t;
        // visit index v is block row v ^ xm
        const int i = e >> 5, k = i + 1 + (e & 31);
        const int ri = i ^ (int)xm, rk = k ^ (int)xm;
        const bool ok = e >= 0 && e < kB * kB && k < kB && pos0 + ri < n && pos0 + rk < n;
        jb_ok[0] |= (uint32_t)ok << e2;
        jb_reg[0][e2] = p.Jt[ok ? (size_t)(pos0 + ri) * n + pos0 + rk : 0];
      }
      const bool hok = hl >= 0 && hl < kB && pos0 + hl < n;
      jb_ok[0] |= (uint32_t)hok << 31;
      h_reg[0] = p.h[hok ? pos0 + hl : 0];
      if (!I8 && p.rowsum) rs_reg = p.rowsum[hok ? pos0 + hl : 0];
    } else {  // each tile's instance (S symmetric: S^T = S; no row sums with int8)
#pragma unroll
      for (int ti = 0; ti < NI; ++ti) {
        const int ni = tile(ti).n;
        const int8_t* sg = p.sgn + tile(ti).sgn;
        jb_ok[ti] = 0;
#pragma unroll
        for (int e2 = 0; e2 < NJB; ++e2) {
          const int e = hl + e2 * NHL;
          const int i = e >> 5, k = i + 1 + (e & 31);
          const int ri = i ^ (int)xm, rk = k ^ (int)xm;
          const bool ok = e >= 0 && e < kB * kB && k < kB && pos0 + ri < ni && pos0 + rk < ni;
          jb_ok[ti] |= (uint32_t)ok << e2;
          jb_reg[ti][e2] = (float)sg[ok ? (size_t)(pos0 + ri) * tile(ti).ld + pos0 + rk : 0];
        }
        const bool hok = hl >= 0 && hl < kB && pos0 + hl < ni;
        jb_ok[ti] |= (uint32_t)hok << 31;
        const int64_t ho = tile(ti).h;
        h_reg[ti] = ho >= 0 ? p.h[ho + (hok ? pos0 + hl : 0)] : 0.f;
      }
    }
  };
  auto publish_block = [&](int pb) {
#pragma unroll
    for (int ti = 0; ti < NI; ++ti)
#pragma unroll
      for (int e2 = 0; e2 < NJB; ++e2) {
        const int e = hl + e2 * NHL;
        if (hl >= 0 && e < kB * kB) L.jb[pb][ti][e] = ((jb_ok[ti] >> e2) & 1u) ? jb_reg[ti][e2] : 0.f;
      }
    if (hl >= 0 && hl < kB) {
#pragma unroll
      for (int ti = 0; ti < NI; ++ti) L.hb[pb][ti][hl] = ((jb_ok[ti] >> 31) & 1u) ? h_reg[ti] : 0.f;
      L.rsb[pb][hl] = (((jb_ok[0] >> 31) & 1u) && !I8 && p.rowsum) ? rs_reg : 0.f;
    }
  };
  // Products of K step j (fragments fa) with this tile's spins hx.
  auto mfma_step = [&](const Frag (&fa)[NST][NP], uint32_t hx, f32x16& acc, i32x16& iacc) {
    if (I8) {
      uint32_t d[4];
      spins_i8(hx, emul, exr, d);
      i32x4 bv;
#pragma unroll
      for (int e = 0; e < 4; ++e) bv[e] = (int)d[e];
      const i32x4 a8 = __builtin_bit_cast(i32x4, fa[0][0]);
      iacc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, bv, iacc, 0, 0, 0);
    } else {
      const uint32_t nz = spins_nz(hx);
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        uint32_t d[4];
        if (st == 0)
          spins_bf16_r<0>(nz, vmask, vone, d);
        else
          spins_bf16_r<1>(nz, vmask, vone, d);
        u32x4v ud = {d[0], d[1], d[2], d[3]};
        const bf16x8 bv = __builtin_bit_cast(bf16x8, ud);
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, fa[st][pc]), bv,
                                                        acc, 0, 0, 0);
      }
    }
  };
  // Fields from the C layout (register v: row 8(v/4) + 4h + v%4 of column
  // lane%32) into [replica][row], four consecutive rows per 16-byte store.
  auto write_fields_col = [&](float* dst, const f32x16& acc, const i32x16& iacc, int qcol) {
#pragma unroll
    for (int g4 = 0; g4 < 4; ++g4) {
      const int i = 8 * g4 + 4 * h;
      float4 x;
      if (I8)
        x = make_float4((float)iacc[4 * g4], (float)iacc[4 * g4 + 1], (float)iacc[4 * g4 + 2],
                        (float)iacc[4 * g4 + 3]);
      else
        x = make_float4(acc[4 * g4], acc[4 * g4 + 1], acc[4 * g4 + 2], acc[4 * g4 + 3]);
      *reinterpret_cast<float4*>(&dst[fld(qcol, i)]) = x;
    }
  };

  // Partial fields of row block cr over this wave's K range, skipping K block
  // ex (its product is added by the proposal wave once the block's spins are
  // final); ring holds the first PF stages of row block cr on entry.
  auto compute_part = [&](int cr, int ex, float* dst) {
    f32x16 acc[NT] = {};
    i32x16 iacc[NT] = {};
    const int jx = ex - c0;  // the skipped stage, if it is in this wave's range
    constexpr int kUnrollJ = NW <= 8 ? (2 * NW + PF - 1) / PF : 1;
#pragma unroll kUnrollJ
    for (int jb0 = 0; jb0 < 2 * NW; jb0 += PF) {
      if (jb0 >= jn) break;
      // stages left from this chunk on, recomputed per chunk (a scalar
      // compare a slot instead of per-slot lane masks kept across the loop)
      int jrem = __builtin_amdgcn_readfirstlane(jn - jb0);
      asm volatile("" : "+s"(jrem));
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int j = jb0 + u;
        // consume slot u, then refill it with stage j + PF (no copy of the
        // fragments: the load is issued after the products that read them);
        // slots past the 2 NW spin halves exist only when PF does not divide
        // 2 NW and are never live
        if (j < 2 * NW && u < jrem && j != jx) {
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
            mfma_step(ring[BAT ? tt : 0][u], get_half<NW>(sa[tt], j), acc[tt], iacc[tt]);
        }
#pragma unroll
        for (int ti = 0; ti < NI; ++ti) load_stage(cr, j + PF, ring[ti][u], ti);
      }
    }
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) write_fields_col(dst, acc[tt], iacc[tt], qtt[tt]);
  };
  // Acceptance brackets of row block cr (in-block mask xm) for sweep counter
  // sw, stored by visit index, lane = replica (beta from L.rb): groups of 4
  // rows (one Philox block each) dealt over the helpers on SIMDs 1..MW, so
  // that none lands on the SIMD of the proposal chain.
  auto brackets = [&](int cr, uint32_t xm, uint64_t sw, int pb) {
    float cb, cz;
    bracket_factors(L.rb[q], cb, cz);
    const int pos0 = 32 * cr;
    for (int g4 = w - MW - 2; g4 < kB / 4; g4 += MW) {
      const u32x4 blk = sweep_block(cid, sw, (uint32_t)(pos0 + 4 * g4), p.k0, p.k1);
      const uint32_t wd[4] = {blk.x, blk.y, blk.z, blk.w};
      float lo[4], hi[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        accept_bounds_f(wd[e], cb, cz, lo[e], hi[e]);
        if (pos0 + 4 * g4 + e >= nq || !act) lo[e] = hi[e] = INFINITY;  // padding, idle lanes
      }
      // rows 4 g4 + e are visits 4 (g4 ^ xm / 4) + (e ^ xm % 4): one 16-byte
      // store each, its entries exchanged by the uniform low bits of xm
      if (xm & 1u) {
        swap2(lo[0], lo[1]), swap2(lo[2], lo[3]);
        swap2(hi[0], hi[1]), swap2(hi[2], hi[3]);
      }
      if (xm & 2u) {
        swap2(lo[0], lo[2]), swap2(lo[1], lo[3]);
        swap2(hi[0], hi[2]), swap2(hi[1], hi[3]);
      }
      const int v4 = 4 * (g4 ^ (int)(xm >> 2));
      *reinterpret_cast<float4*>(&L.lo[pb][q][v4]) = make_float4(lo[0], lo[1], lo[2], lo[3]);
      *reinterpret_cast<float4*>(&L.hi[pb][q][v4]) = make_float4(hi[0], hi[1], hi[2], hi[3]);
    }
  };
  auto owner_of = [&](int c) { return 1 + c / kb; };  // matrix wave owning K block c
  // by the owner of block c, in the visit order of its next visit (mask xm),
  // so that the proposal chain starts without a permute of its own
  auto publish_bits = [&](int c, int pb, uint32_t xm) {
    const int j = c - c0;
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
      const uint32_t mine = get_half<NW>(sa[tt], j);
      const uint32_t got = (uint32_t)__shfl_xor((int)mine, 32, 64);  // the other half
      const uint32_t bits = block_of<I8>(mine, got);
      if (lane < 32) L.bb[pb][qtt[tt]] = BR ? xor_permute_bits(bits, xm) : bits;
    }
  };

  // Visit schedule of sweep counter swn into parity par (wave 4, one lane
  // per step): the block visited at step t and that block's in-block mask.
  auto fill_sched = [&](int par, uint64_t swn, bool dry) {
    if (lane < nblk) {
      int bo = lane;
      uint32_t mk = 0;
      if (p.order == kBlockRandom && !dry) {
        const FeistelKey fk = block_order_key((uint32_t)nblk, swn, p.k0, p.k1);
        bo = (int)feistel_perm(fk, (uint32_t)lane);
        mk = block_mask(swn, (uint32_t)bo, p.k0, p.k1);
      }
      L.sch[par][lane] = (uint16_t)(bo | (mk << 8));
    }
  };
  // The entry of a step: block (low byte) and mask, one LDS read for both.
  // Reads are volatile so that each use reloads the entry (a short LDS read)
  // instead of keeping it live in scalar registers across the block's work.
  auto sched_entry = [&](int par, int step) -> uint32_t {
    if (!BR) return (uint32_t)step;
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)*(volatile uint16_t*)&L.sch[par][step]);
  };
  auto entry_block = [](uint32_t e) { return BR ? (int)(e & 0xffu) : (int)e; };
  auto entry_mask = [](uint32_t e) { return BR ? e >> 8 : 0u; };

  // Passes over the blocks.  Every `refresh` sweeps (p.refresh > 0 and
  // S > refresh) a dry pass recomputes the energy: the tracked energy adds
  // each sweep's accepted changes (simulation.py:157-158), and each change
  // carries the f32 rounding of the matrix-product fields, which
  // random-walks over millions of flips (the reference recomputes each field
  // from its coupling row).  A dry pass runs the pipeline over every block
  // with the proposals switched off, so each block's fields are the full
  // products F = J s of the current state (with the couplings the products
  // use); wave 0 sums E = s.(F/2 + h) in f64 and resets the tracked energy
  // to it plus the offset the caller's initial energy had from the dry pass
  // before the first sweep (zero unless the caller passed other energies).
  // Integer couplings are exact either way.  Passes: a dry pass, then
  // `refresh` sweeps and a dry pass, repeated; no dry pass after the last
  // sweep.  pass_info gives a pass's kind and its sweep (for a dry pass the
  // sweep that follows it).
  const int kre = p.refresh;
  const bool rec = kre > 0 && p.S > kre;
  const int npass = rec ? p.S + 1 + (p.S - 1) / kre : p.S;
  auto pass_info = [&](int ps, int& s_out) -> bool {
    if (!rec) {
      s_out = ps;
      return false;
    }
    if (ps == 0) {
      s_out = 0;
      return true;
    }
    const int qq = ps - 1, bq = qq / (kre + 1), rq = qq - bq * (kre + 1);
    if (rq == kre) {
      s_out = (bq + 1) * kre;
      return true;
    }
    s_out = bq * kre + rq;
    return false;
  };

  // ---- prologue: the fields of the first block in full, its brackets and spins
  {
    int s0;
    const bool dry0 = pass_info(0, s0);
    if (BR && w == 4) fill_sched(0, (uint64_t)s0 + p.sweep_offset, dry0);
  }
  __syncthreads();  // wave 0's betas (L.rb), the zeroed L.corr and the schedule are visible
  {
    int s0;
    (void)pass_info(0, s0);
    const uint64_t sw0 = (uint64_t)s0 + p.sweep_offset;
    const uint32_t ef = sched_entry(0, 0), es = sched_entry(0, 1);
    const int cf = entry_block(ef), cs = entry_block(es);
    const uint32_t mf = entry_mask(ef), ms = entry_mask(es);
    if (mw) {
      prefetch_frags(cf);
      compute_part(cf, -1, L.part[0][m]);
      prefetch_frags(cs);
      if (w == owner_of(cf)) publish_bits(cf, 0, mf);
    }
    if (hw) {
      prefetch_jb(cf, mf);
      publish_block(0);
      prefetch_jb(cs, ms);
      if (w > MW + 1) brackets(cf, mf, sw0, 0);
    }
  }
  __syncthreads();

  const float cu = (p.ptype == 0) ? 2.f : 1.f;  // |dn / (F + h)|
  // F = fsc * (sum of the partial products + rsb) + h: QUBO with bf16
  // operands used +-1 spins, F = (J.s + J.1) / 2 with the row sums J.1 in
  // rsb (zero otherwise); fsc is a power of two, so the fma rounds like the add
  const bool qsum = !I8 && p.ptype == 1;
  const float fsc = qsum ? 0.5f : 1.f;
  // batched: field = bsc * count + h, the scale of the lane's instance
  const float bsc = BAT ? (tq ? tile(1).scale : tile(0).scale) : 1.f;
  double eoff = 0.0;  // wave 0: caller's initial energy minus its recomputation

  // The correction J[c1, c] . S_c of the next block c1 by the proposal wave,
  // once block c's spins are final (on SIMD 0's matrix core, idle otherwise):
  // bits = block c's spins of replica `lane` in block-row order; the A
  // fragments of row block c1, K block c were loaded into fa at the block's
  // start.  Result to L.corr (read by this wave at the next block's start).
  typedef Frag FragSet[NI][NST][NP];  // by instance
  auto load_corr_frags = [&](int c1, int c, FragSet& fa) {
#pragma unroll
    for (int ti = 0; ti < NI; ++ti) {
      const Frag* fb = frag_base(ti);
#pragma unroll
      for (int st = 0; st < NST; ++st)
#pragma unroll
        for (int pc = 0; pc < NP; ++pc) {
          const size_t f = I8 ? (size_t)c1 * nblk + c
                              : ((size_t)c1 * (2 * nblk) + 2 * c + st) * NP + pc;
          fa[ti][st][pc] = fb[f * 64 + lane];
        }
    }
  };
  auto corr_products = [&](const FragSet& fa, uint32_t bits, int ql) {  // ql: lane
#pragma unroll
    for (int tl = 0; tl < NT; ++tl) {
      // column lane % 32 of tile tl is replica 32 tl + lane % 32
      const uint32_t bt = (uint32_t)__shfl((int)bits, 32 * tl + (ql & 31), 64);
      f32x16 a = {};
      i32x16 ia = {};
      mfma_step(fa[BAT ? tl : 0], half_of<I8>(bt, h), a, ia);
      write_fields_col(L.corr, a, ia, 32 * tl + (ql & 31));
    }
  };

  uint32_t t = 0;                               // running block counter
  // The two roles run separate copies of the block loop (so that each keeps
  // only its own state in registers); both pass the same barriers: one after
  // the block's work (B), and at the end of a pass one more (three after a
  // sweep).
  if (w == 0) {
    // schedule entries of the current step, each read one block ahead
    const uint32_t e0 = sched_entry(0, 0);
    int c_nx = entry_block(e0);
    uint32_t xm_nx = entry_mask(e0);
    for (int ps = 0; ps < npass; ++ps) {
      int s;
      const bool dry = pass_info(ps, s);
      const uint64_t sw = (uint64_t)s + p.sweep_offset;
      const int sp = ps & 1;
      const float beta = act ? L.rb[q] : 0.f;
      double dsum = 0.0, eacc = 0.0;
      uint32_t nacc = 0;
      for (int tb = 0; tb < nblk; ++tb, ++t) {
        const int pb = t & 1;
        // The lane's LDS offsets are recomputed each block from an opaque
        // copy of its index (a few VALU ops): hoisted out of the loop they
        // would stay live across the chain, where registers are scarce.
        int qo = q;
        asm volatile("" : "+v"(qo));
        // block c of this step, its visit k being block row k ^ xm; c1 the next
        const int c = c_nx;
        const uint32_t xm = xm_nx;
        const int pos0 = 32 * c;
        const bool lastb = tb + 1 == nblk;
        const uint32_t e1 = sched_entry(lastb ? sp ^ 1 : sp, lastb ? 0 : tb + 1);
        const int c1 = entry_block(e1);
        xm_nx = entry_mask(e1);
        c_nx = c1;
        FragSet fcorr;
        load_corr_frags(c1, c, fcorr);
        if (dry) {
          // ---- dry pass: f64 energy of the block's rows from its full fields
          // (a dry pass visits every block with mask 0: bits in row order)
          const uint32_t bits = L.bb[pb][qo];
          for (int i = 0; i < kB; ++i) {
            if (pos0 + i < nq) {
              const int o = fld(qo, i);
              double f = (double)L.part[pb][0][o];
              for (int mm = 1; mm < MW; ++mm) f += (double)L.part[pb][mm][o];
              f += (double)L.corr[o];
              if (qsum) f = 0.5 * (f + (double)L.rsb[pb][i]);
              if (BAT) f *= (double)bsc;
              const double si = ((bits >> i) & 1u) ? 1.0 : (p.ptype == 0 ? -1.0 : 0.0);
              eacc += si * (0.5 * f + (double)L.hb[pb][tq][i]);
            }
          }
          L.bnew[pb][qo] = bits;
          // the same block next (the owners skip it), in its next visit order
          if (c1 == c) L.bb[pb ^ 1][qo] = BR ? xor_permute_bits(bits, xm_nx) : bits;
          corr_products(fcorr, bits, qo);
          __syncthreads();  // B
          if (lastb) {
            if (act) {
              if constexpr (AIS) {
                if (ps == 0)
                  eoff = Ed - eacc;
                else
                  Ed = eacc + eoff;
              } else {
                if (ps == 0)
                  eoff = (double)E - eacc;
                else
                  E = (float)(eacc + eoff);
              }
            }
            __syncthreads();  // the next pass's first brackets
          }
          continue;
        }
        // ---- proposals of block c, one lane per replica
        // brackets of the block's 32 visits, read first (their latency
        // overlaps the field sums and they leave the chain free of LDS waits)
        float hib[kB], lob[kB];
#pragma unroll
        for (int t4 = 0; t4 < kB / 4; ++t4) {
          const float4 a = *reinterpret_cast<const float4*>(&L.hi[pb][qo][4 * t4]);
          const float4 b = *reinterpret_cast<const float4*>(&L.lo[pb][qo][4 * t4]);
          hib[4 * t4] = a.x, hib[4 * t4 + 1] = a.y, hib[4 * t4 + 2] = a.z, hib[4 * t4 + 3] = a.w;
          lob[4 * t4] = b.x, lob[4 * t4 + 1] = b.y, lob[4 * t4 + 2] = b.z, lob[4 * t4 + 3] = b.w;
        }
        // fields in visit order (batched: the integer counts, with h in hk)
        float fp[kB], hk[kB];
        {
          const int xg = (int)(xm >> 2);
          // visit group r4 holds block rows 4 (r4 ^ xg) + (e ^ (xm & 3)); the
          // in-group part is a compile-time permutation per uniform branch
          // FM: 0 = sums + h; 1 = (sums + row sums) / 2 + h (QUBO with bf16
          // operands); 2 = the sums alone (every h zero: adding +0 would only
          // turn a -0 into +0, which no decision or sum distinguishes)
          auto load_fields = [&](auto pc, auto fmc) __attribute__((always_inline)) {
            constexpr int P = decltype(pc)::value;
            constexpr int FM = decltype(fmc)::value;
            constexpr bool RS = FM == 1;
            // kFG groups of 4 rows per stage: every LDS read of a stage is
            // issued before its first sum, so a stage costs one LDS latency
#pragma unroll
            for (int g0 = 0; g0 < kB / 4; g0 += kFG) {
              float4 ld[kFG][MW + 3];
#pragma unroll
              for (int j = 0; j < kFG; ++j) {
                const int gr = (g0 + j) ^ xg;
                const int o = fld(qo, 4 * gr);
#pragma unroll
                for (int mm = 0; mm < MW; ++mm)
                  ld[j][mm] = *reinterpret_cast<const float4*>(&L.part[pb][mm][o]);
                ld[j][MW] = *reinterpret_cast<const float4*>(&L.corr[o]);
                if constexpr (FM != 2) ld[j][MW + 1] = *reinterpret_cast<const float4*>(&L.hb[pb][tq][4 * gr]);
                if constexpr (RS) ld[j][MW + 2] = *reinterpret_cast<const float4*>(&L.rsb[pb][4 * gr]);
              }
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int j = 0; j < kFG; ++j) {
                float4 sm = ld[j][0];
#pragma unroll
                for (int mm = 1; mm < MW; ++mm) {
                  sm.x = sm.x + ld[j][mm].x;
                  sm.y = sm.y + ld[j][mm].y;
                  sm.z = sm.z + ld[j][mm].z;
                  sm.w = sm.w + ld[j][mm].w;
                }
                const float4 cr = ld[j][MW];
                const float4 h4 = FM != 2 ? ld[j][MW + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
                float vs[4] = {sm.x + cr.x, sm.y + cr.y, sm.z + cr.z, sm.w + cr.w};
                if constexpr (RS) {
                  const float4 s4 = ld[j][MW + 2];
                  vs[0] = vs[0] + s4.x, vs[1] = vs[1] + s4.y, vs[2] = vs[2] + s4.z, vs[3] = vs[3] + s4.w;
                }
                const float hv[4] = {h4.x, h4.y, h4.z, h4.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  if constexpr (BAT) {
                    fp[4 * (g0 + j) + e] = vs[e ^ P];
                    hk[4 * (g0 + j) + e] = hv[e ^ P];
                  } else {
                    if constexpr (FM == 2)
                      fp[4 * (g0 + j) + e] = vs[e ^ P];
                    else
                      fp[4 * (g0 + j) + e] = __builtin_fmaf(fsc, vs[e ^ P], hv[e ^ P]);
                  }
                }
              }
              __builtin_amdgcn_sched_barrier(0);
            }
          };
          auto load_fm = [&](auto fmc) __attribute__((always_inline)) {
            switch (xm & 3u) {
              case 0: load_fields(std::integral_constant<int, 0>{}, fmc); break;
              case 1: load_fields(std::integral_constant<int, 1>{}, fmc); break;
              case 2: load_fields(std::integral_constant<int, 2>{}, fmc); break;
              default: load_fields(std::integral_constant<int, 3>{}, fmc); break;
            }
          };
          if (qsum)
            load_fm(std::integral_constant<int, 1>{});
          else if (!BAT && p.hzero)
            load_fm(std::integral_constant<int, 2>{});
          else
            load_fm(std::integral_constant<int, 0>{});

        }
        // spins in visit order (bit k = block row k ^ xm), as published
        uint32_t mb = L.bb[pb][qo];
        const uint32_t mb0 = mb;
        // accepted flips, step k at bit 31 - k: one add with carry per step
        uint32_t tk = 0;
        // bit k of mb is only changed at step k, so every step reads its
        // spin from the block's initial bits: sign bit = !spin, magnitude cu
        const uint32_t nmb = ~mb0, cub = __float_as_uint(cu);
        uint32_t vsgn = 0x80000000u;  // kept in a register (not a literal operand)
        asm volatile("" : "+v"(vsgn));
        float dpart = 0.f;
        const double dsum0 = dsum;  // batched: the sum before the block (for a rerun)
        uint64_t amb = 0;  // lanes with a decision inside its bracket
        // Rows of J[blk][blk] (the same for every lane) stream from LDS
        // through a ring JR rows ahead of the chain.  Row k holds only the
        // columns right of the diagonal, shifted to the start of the row, so
        // the ceil((31 - k) / 4) 16-byte loads of a row carry no dead lanes.
        constexpr int JR = 2;
        const float4* jb4 = reinterpret_cast<const float4*>(L.jb[pb][tq]);
        float4 jring[JR][kB / 4];
        static_for<JR>([&](auto dc) __attribute__((always_inline)) {
          constexpr int d = decltype(dc)::value;
          static_for<(kB - 1 - d + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
            constexpr int t4 = decltype(tc)::value;
            jring[d][t4] = jb4[d * (kB / 4) + t4];
          });
        });
        static_for<kB>([&](auto kc) __attribute__((always_inline)) {
          constexpr int k = decltype(kc)::value;
          // dn = 2 s (F + h) (ising) or (2x - 1)(F + h) (qubo), exact in f32
          const uint32_t shk = nmb << (31 - k);
          const float sg = __uint_as_float((shk & vsgn) | (cub & ~vsgn));  // one bit insert
          // batched: the field scale * c + h of the count (a product, then a sum)
          const float fk = BAT ? fp[k] * bsc + hk[k] : fp[k];
          const float dn = fk * sg;
          const float hik = hib[k], lok = lob[k];
          // idle lanes and padding rows have lo = hi = +inf (and finite dn):
          // never taken, never inside their bracket
          const bool take = dn > hik;
          // the decision's lane mask, used as it is by the two selects below
          const uint64_t tm = __builtin_amdgcn_ballot_w64(take);
          amb |= __builtin_amdgcn_ballot_w64(!(dn < lok)) & ~tm;
          float dS;  // take ? -sg : 0
          asm("v_cndmask_b32_e64 %0, 0, -%1, %2" : "=v"(dS) : "v"(sg), "s"(tm));
          {
            uint64_t co;  // tk = 2 tk + take (the mask as carry-in)
            asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(tk), "=s"(co) : "v"(tk), "s"(tm));
            (void)co;
          }
          // dS * fk is exactly -dn or 0, so this rounds like dpart - dn
          // (batched: summed in f64 in visit order, as the oracle sums)
          if constexpr (BAT)
            dsum += (double)(dS * fk);
          else
            dpart = __builtin_fmaf(dS, fp[k], dpart);
          static_for<(kB - 1 - k + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
            constexpr int t4 = decltype(tc)::value;
            const float4 j4 = jring[k % JR][t4];
            const float jv[4] = {j4.x, j4.y, j4.z, j4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int r = k + 1 + 4 * t4 + e;
              // dS is 0, +-1 or +-2, so the fused form rounds like mul + add
              if (r < kB) fp[r] = __builtin_fmaf(jv[e], dS, fp[r]);
            }
          });
          if (k + JR < kB)
            static_for<(kB - 1 - (k + JR) + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
              constexpr int t4 = decltype(tc)::value;
              jring[k % JR][t4] = jb4[(k + JR) * (kB / 4) + t4];
            });
        });
        mb = mb0 ^ __builtin_bitreverse32(tk);
        if (__builtin_expect(amb != 0, 0)) {
          // A decision fell inside its bracket: rerun the block, taking the
          // exact test for the decisions inside their brackets (rolled loop,
          // fields staged in LDS).
          float* sf = &L.sfp[qo * 33];
          for (int r = 0; r < kB; ++r) {  // visit order, as fp
            const int br = r ^ (int)xm;
            float v = L.part[pb][0][fld(qo, br)];
            for (int mm = 1; mm < MW; ++mm) v = v + L.part[pb][mm][fld(qo, br)];
            v = v + L.corr[fld(qo, br)];
            sf[r] = BAT ? v : __builtin_fmaf(fsc, v + L.rsb[pb][br], L.hb[pb][tq][br]);
          }
          mb = mb0;
          dpart = 0.f;
          if constexpr (BAT) dsum = dsum0;
          for (int k = 0; k < kB; ++k) {
            const bool up = (mb >> k) & 1u;
            const float sg = up ? cu : -cu;
            const float fk = BAT ? sf[k] * bsc + L.hb[pb][tq][k ^ (int)xm] : sf[k];
            const float dn = fk * sg;
            // idle lanes and padding rows have lo = hi = +inf: never taken
            bool take = act & (dn > L.hi[pb][qo][k]);
            if (act && !take && !(dn < L.lo[pb][qo][k])) {
              const uint32_t pk = (uint32_t)(pos0 + (k ^ (int)xm));
              const uint32_t wd = pick_word(sweep_block(cid, sw, pk, p.k0, p.k1), pk & 3u);
              take = accept_exact(wd, beta, dn);
            }
            mb ^= (uint32_t)take << k;
            if constexpr (BAT)
              dsum += take ? (double)(-dn) : 0.0;
            else
              dpart += take ? -dn : 0.f;
            const float dS = take ? -sg : 0.f;
            for (int r = k + 1; r < kB; ++r)
              sf[r] = __builtin_fmaf(L.jb[pb][tq][k * kB + (r - k - 1)], dS, sf[r]);
          }
        }
        {
          const uint32_t mbr = BR ? xor_permute_bits(mb, xm) : mb;  // back to block-row order
          L.bnew[pb][qo] = mbr;
          // the same block next (at a pass boundary): its spins are these,
          // in its next visit order (the owners, whose copy is older, skip it)
          if (c1 == c) L.bb[pb ^ 1][qo] = BR ? xor_permute_bits(mbr, xm_nx) : mbr;
          corr_products(fcorr, mbr, qo);
        }
        if constexpr (!BAT) dsum += (double)dpart;
        nacc += (uint32_t)__builtin_popcount(mb ^ mb0);
        __syncthreads();  // B
        if (AIS && lastb) {
          // ---- end of an AIS step (simulation.py:305-316 with the uniform
          // reference: its log-density is 0): the energy after the sweep, the
          // weight of this step (ordering 0) or of the next one before its
          // sweep (ordering 1), and the next step's beta for every chain
          acc += nacc;
          Ed = Ed + dsum;
          const double bt = (double)p.temps[s];
          if (act && p.ordering == 0) lw = lw + ((-bt * Ed) - (-bold * Ed));
          bold = bt;
          const bool more = s + 1 < p.S;
          const float tn = more ? p.temps[s + 1] : 0.f;
          if (act && p.ordering == 1 && more) lw = lw + ((-(double)tn * Ed) - (-bold * Ed));
          L.rb[q] = act ? tn : 0.f;
          __syncthreads();
          __syncthreads();
          __syncthreads();  // the next pass's first brackets
        } else if (lastb) {
          // ---- end of sweep: energy (simulation.py:157-158), best states, PT
          acc += nacc;
          E = (float)((double)E + dsum);
          L.E[q] = E;
          if (p.best_sem == 0) {  // simulation.py:161-164
            const bool imp = act && (E < bestE);
            if (imp) {
              bestE = E;
              p.best_rep_e[(size_t)read * R + rr] = E;
              p.best_rep_sw[(size_t)read * R + rr] = s;
            }
            L.flag[q] = imp ? 1 : 0;
          }
          __syncthreads();
          // lane b0 + k: the energy of the replica in slot k (the walk below
          // reads the slots in order; the replicas' energies do not change).
          // The lane offsets are derived here from an opaque copy of the lane
          // index: hoisted to the kernel start they were kept in scratch.
          int qx = q;
          asm volatile("" : "+v"(qx));
          const int tx = BAT ? (qx >> 5) : 0;
          const int b0 = BAT ? 32 * tx + ((qx & 31) / R) * R : (qx / R) * R;  // the read's first lane
          const int slot = qx - b0;
          const float* bet = &L.bet[32 * tx];
          if (act) L.Es[qx] = L.E[b0 + L.bi[qx]];
          if (act && rr == 0) {  // read leader
            int src = -1;
            if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
              float cur = gE;
              float en = L.E[b0];
              for (int k = 0; k < R; ++k) {
                const float ek = en;
                if (k + 1 < R) en = L.E[b0 + k + 1];  // the next one, read ahead
                if (ek < cur) {
                  cur = ek;
                  src = k;
                }
              }
              if (src >= 0) {
                gE = cur;
                gsw = s;
                gimp = 1;
              }
            }
            if (p.use_pt && R > 1) {  // utils.py:44-58
              // Hot end first: slot a = R-k-1 holds the carried replica (the one
              // the previous pair left there), slot b = a-1 still its occupant
              // from before the walk.  An exchange moves the carried replica on
              // to b; otherwise b's occupant becomes the carried one.  Every
              // operand but the carried one is read one step ahead.
              int carry = L.bi[b0 + R - 1];
              float ec = L.Es[b0 + R - 1], ba = bet[R - 1];
              int rbn = L.bi[b0 + R - 2];
              float ebn = L.Es[b0 + R - 2], bbn = bet[R - 2];
              double lun = L.lu[b0];
              for (int k = 0; k < R - 1; ++k) {
                const int a = R - k - 1, b = a - 1;
                const int rb = rbn;
                const float eb = ebn, bb = bbn;
                const double luk = lun;
                if (b > 0) {
                  rbn = L.bi[b0 + b - 1];
                  ebn = L.Es[b0 + b - 1];
                  bbn = bet[b - 1];
                  lun = L.lu[b0 + k + 1];
                }
                const float dd = (ec - eb) * (ba - bb);
                const bool x = dd >= 0.f || luk < (double)dd;
                L.bi[b0 + a] = x ? rb : carry;
                if (!x) {
                  carry = rb;
                  ec = eb;
                }
                ba = bb;
              }
              L.bi[b0] = carry;
            }
            if (p.logw && p.log_omegas) {  // ais.py:99-112 over the ascending ladder
              double lw = 0.0, bold = 0.0;
              for (int r = 1; r < R; ++r) {
                const int k = L.bso[r];
                const float be = bet[k];
                const float en = L.E[b0 + L.bi[b0 + k]];
                const float a = -be * en;
                lw += (double)a - (-bold * (double)en);
                bold = (double)be;
              }
              p.log_omegas[(size_t)read * p.S + s] = lw;
            }
            L.src[g] = src;
          }
          // every slot's new beta to the replica now in it (after the
          // leader's exchanges: the same wave, in order)
          if (act) L.rb[b0 + L.bi[qx]] = bet[slot];
          __syncthreads();
          __syncthreads();  // the next pass's first brackets
        }
      }
    }
  } else {
    // schedule entries carried from block to block: c (this step), c1 (the
    // next) and its mask; c2 (the one after) is read each block
    const uint32_t e1 = sched_entry(0, 1);
    int c = entry_block(sched_entry(0, 0)), c1 = entry_block(e1);
    uint32_t m1 = entry_mask(e1);
    for (int ps = 0; ps < npass; ++ps) {
      int s, s_next;
      const bool dry = pass_info(ps, s);
      const bool dry_next = pass_info(ps + 1, s_next);
      const uint64_t sw = (uint64_t)s + p.sweep_offset;
      const uint64_t sw_next = (uint64_t)s_next + p.sweep_offset;
      const int sp = ps & 1;
      for (int tb = 0; tb < nblk; ++tb, ++t) {
        const int pb = t & 1, pn = pb ^ 1;
        // the schedule of the next pass, needed from step nblk - 2 on
        if (BR && w == 4 && tb == 0) fill_sched(sp ^ 1, sw_next, dry_next);
        const bool last = tb + 1 == nblk;
        const int p2 = tb + 2 < nblk ? sp : sp ^ 1, i2 = tb + 2 < nblk ? tb + 2 : tb + 2 - nblk;
        // The entry two steps ahead is read here and made wave-uniform where
        // it is first needed (the prefetches at the end of the step), so the
        // step's work does not wait for the LDS read.
        const uint32_t raw2 = BR ? (uint32_t)*(volatile uint16_t*)&L.sch[p2][i2] : (uint32_t)i2;
        auto entry2 = [&]() {
          return BR ? (uint32_t)__builtin_amdgcn_readfirstlane((int)raw2) : raw2;
        };
        int c2 = 0;
        uint32_t m2 = 0;
        // ---- meanwhile: block c1's spins, fields except K block c, J / h and brackets
        if (mw) {
          // c1's spins have not changed since its last visit unless c1 == c
          // (a pass boundary), which the proposal wave publishes itself
          if (w == owner_of(c1) && c1 != c) publish_bits(c1, pn, m1);
          compute_part(c1, c, L.part[pn][m]);
          const uint32_t en2 = entry2();
          c2 = entry_block(en2);
          m2 = entry_mask(en2);
          prefetch_frags(c2);
        }
        if (hw) {
          publish_block(pn);
          if (w > MW + 1 && !last && !dry) brackets(c1, m1, sw, pn);
          const uint32_t en2 = entry2();
          c2 = entry_block(en2);
          m2 = entry_mask(en2);
          prefetch_jb(c2, m2);
          // the sweep's exchange uniforms (utils.py:57), by a helper that has
          // no brackets to compute during the last block
          if (w == MW + 2 && last && !dry && p.use_pt && act && rr < R - 1)
            L.lu[q] = log_u32(pick_word(pt_block((uint64_t)read, sw, (uint32_t)rr, p.k0, p.k1),
                                        (uint32_t)rr & 3u));
        }
        __syncthreads();  // B
        // ---- the owners of block c take its new spins
        if (mw && w == owner_of(c)) {
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
            set_half<NW>(sa[tt], c - c0, half_of<I8>(L.bnew[pb][qtt[tt]], h));
        }
        if (last) {
          if (!dry) {
            __syncthreads();
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
              if (!AIS && mw && p.best_sem == 0 && actt[tt] && L.flag[qtt[tt]])
                store_range<NW, I8>(sa[tt], c0, kb, nblk, n_of(tt), p.ptype,
                                    p.best_rep + base_t[tt]);
            __syncthreads();
            // owners copy their part of a new global best
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
              if (!AIS && mw && p.best_sem == 1 && actt[tt] && L.src[gtt[tt]] == rtt_of(tt))
                store_range<NW, I8>(sa[tt], c0, kb, nblk, n_of(tt), p.ptype, best_of(tt));
          }
          // the next pass's first block (after the exchanges: new betas)
          if (hw && w > MW + 1 && !dry_next) brackets(c1, m1, sw_next, pn);
          __syncthreads();
        }
        c = c1;
        c1 = c2;
        m1 = m2;
      }
    }
  }

  // ---- outputs
#pragma unroll
  for (int tt = 0; tt < NT; ++tt)
    if (actt[tt]) store_range<NW, I8>(sa[tt], c0, kb, nblk, n_of(tt), p.ptype, p.spins + base_t[tt]);
  if (AIS && w == 0 && act) {
    if (p.accepted) atomicAdd(p.accepted + read, acc);
    p.ais_e[read] = Ed;
    p.ais_w[read] = lw;
  }
  if (!AIS && w == 0 && act) {
    if (p.accepted) atomicAdd(p.accepted + read, acc);
    p.energies[(size_t)read * R + rr] = E;
    p.beta_idx[(size_t)read * R + rr] = L.bi[q];
    if (rr == 0 && p.best_sem == 1) {
      p.best_energy[read] = gE;
      p.best_sweeps[read] = gsw;
      L.src[g] = gimp;
    }
  }
  if (!AIS && p.best_sem == 1) {
    // never improved: best_state is still the view of states[0] (simulation.py:204)
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
      if (actt[tt] && rtt_of(tt) == 0 && !L.src[gtt[tt]])
        store_range<NW, I8>(sa[tt], c0, kb, nblk, n_of(tt), p.ptype, best_of(tt));
  }
}

template <int NP, bool I8, bool BR, int NT, int MODE = kModePT>
hipError_t launch_pipe_o(const BParams& p, int nw, int groups, hipStream_t st) {
  dim3 grid(groups), block(128 * (kPipeMW + 1));
  switch (nw) {
    case 4: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 4, kPipeMW, BR, NT, MODE>), grid, block, 0, st, p); break;
    case 8: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 8, kPipeMW, BR, NT, MODE>), grid, block, 0, st, p); break;
    case 16: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 16, kPipeMW, BR, NT, MODE>), grid, block, 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Launch of one coupling encoding: nw spin words per matrix wave, nt tiles of
// 32 replicas per workgroup, the order from p.
template <int NP, bool I8>
hipError_t launch_pipe_enc(const BParams& p, int nw, int nt, int groups, hipStream_t st) {
  if (nt != 1 && nt != 2) return hipErrorInvalidValue;
  if (p.order == kBlockRandom)
    return nt == 2 ? launch_pipe_o<NP, I8, true, 2>(p, nw, groups, st)
                   : launch_pipe_o<NP, I8, true, 1>(p, nw, groups, st);
  return nt == 2 ? launch_pipe_o<NP, I8, false, 2>(p, nw, groups, st)
                 : launch_pipe_o<NP, I8, false, 1>(p, nw, groups, st);
}

// AIS (kModeAIS) with one coupling encoding: two tiles of 32 chains per
// workgroup, the order from p.
template <int NP, bool I8>
hipError_t launch_pipe_ais_enc(const BParams& p, int nw, int groups, hipStream_t st) {
  return p.order == kBlockRandom ? launch_pipe_o<NP, I8, true, 2, kModeAIS>(p, nw, groups, st)
                                 : launch_pipe_o<NP, I8, false, 2, kModeAIS>(p, nw, groups, st);
}

// One per translation unit (blocked_pipe_i8.hip, blocked_pipe_bf16_{1,2,3}.hip).
hipError_t launch_pipe_i8(const BParams& p, int nw, int nt, int groups, hipStream_t st);
hipError_t launch_pipe_bf16_1(const BParams& p, int nw, int nt, int groups, hipStream_t st);
hipError_t launch_pipe_bf16_2(const BParams& p, int nw, int nt, int groups, hipStream_t st);
hipError_t launch_pipe_bf16_3(const BParams& p, int nw, int nt, int groups, hipStream_t st);
hipError_t launch_pipe_ais_i8(const BParams& p, int nw, int groups, hipStream_t st);
hipError_t launch_pipe_ais_bf16_1(const BParams& p, int nw, int groups, hipStream_t st);
hipError_t launch_pipe_ais_bf16_2(const BParams& p, int nw, int groups, hipStream_t st);
hipError_t launch_pipe_ais_bf16_3(const BParams& p, int nw, int groups, hipStream_t st);
// The batched form (blocked_pipe_bat.hip): one workgroup per entry of p.bat.
hipError_t launch_pipe_bat(const BParams& p, int nw, int groups, hipStream_t st);

}  // namespace blk
}  // namespace pysa
