// Batched MFMA form: the pipelined kernel over many instances of different
// sizes in one launch, two tiles of up to 32 chains per workgroup (its entry
// of a BatWG table; the tiles' instances have the same number of 32-spin
// blocks), with the instances' int8 coupling matrices packed into MFMA
// fragments at the start of the call.
//
// PySA has no batched path: a batch is a caller's loop over Solver objects,
// each running pysa/sa.py:92-430 with the reads one after another
// (pysa/sa.py:425-427).  Per instance the chain is the one of
// pysa/simulation.py:123-236 in sequential or block-random order, with the
// couplings J = scale * S of the batched form (S int8) and the field of a
// spin the f32 value scale * (S s) + h.
#include "blocked_pipe.hpp"

namespace pysa {
namespace blk {

hipError_t launch_pipe_bat(const BParams& p, int nw, int groups, hipStream_t st) {
  return p.order == kBlockRandom ? launch_pipe_o<1, true, true, 2, kModeBatched>(p, nw, groups, st)
                                 : launch_pipe_o<1, true, false, 2, kModeBatched>(p, nw, groups, st);
}

namespace {

// Fragment (c, step) of instance y's int8 encoding: lane l holds
// S[32c + l%32][32 step + 16 (l/32) + e], e < 16 (zero outside the matrix).
// Rows are ld bytes (a multiple of 16, zero beyond n), so a lane's 16 bytes
// are one aligned load or all zero.
__global__ void pack_i8_batched_kernel(const int8_t* __restrict__ signs,
                                       const BatPack* __restrict__ tab,
                                       int8_t* __restrict__ out) {
  const BatPack t = tab[blockIdx.y];
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)t.nblk * t.nblk * 64) return;
  const int lane = (int)(gid & 63);
  const long long f = gid >> 6;
  const int step = (int)(f % t.nblk), c = (int)(f / t.nblk);
  const int row = 32 * c + (lane & 31);
  const int col0 = 32 * step + 16 * (lane >> 5);
  int4 v = make_int4(0, 0, 0, 0);
  if (row < t.n && col0 < t.ld)
    v = *reinterpret_cast<const int4*>(signs + t.sgn + (long long)row * t.ld + col0);
  *reinterpret_cast<int4*>(out + t.frag + gid * 16) = v;
}

// Per-read minimum over the replicas' bests (simulation.py:174-177, first
// minimum wins), one thread per read slot of a workgroup's tiles.
__global__ void bat_best_argmin_kernel(BParams p, int groups) {
  const int gw = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = 2 * p.rpw;
  const int wg = gw / per, tt = (gw - wg * per) / p.rpw, j = gw - wg * per - tt * p.rpw;
  if (wg >= groups) return;
  const BatTile& b = p.bat[wg].t[tt];
  const int read = b.read0 + j;
  if (read >= b.read_end) return;
  const int R = p.R, n = b.n;
  int bp = 0;
  for (int k = 1; k < R; ++k)
    if (p.best_rep_e[(size_t)read * R + k] < p.best_rep_e[(size_t)read * R + bp]) bp = k;
  p.best_energy[read] = p.best_rep_e[(size_t)read * R + bp];
  p.best_sweeps[read] = p.best_rep_sw[(size_t)read * R + bp];
  const int8_t* src = p.best_rep + (b.spin_adj + ((int64_t)read * R + bp) * n);
  int8_t* dst = p.best_state + (b.best_adj + (int64_t)read * n);
  for (int i = 0; i < n; ++i) dst[i] = src[i];
}

}  // namespace

hipError_t pack_batched_frags(const int8_t* signs, const BatPack* tab, int count, int max_nblk,
                              int8_t* out, hipStream_t st) {
  const long long lanes = (long long)max_nblk * max_nblk * 64;
  for (int i0 = 0; i0 < count; i0 += 65535) {  // grid.y stays within its limit
    const int cnt = count - i0 < 65535 ? count - i0 : 65535;
    const dim3 grid((unsigned)((lanes + 255) / 256), (unsigned)cnt), block(256);
    hipLaunchKernelGGL(pack_i8_batched_kernel, grid, block, 0, st, signs, tab + i0, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t bat_best_argmin(const BParams& p, int groups, hipStream_t st) {
  const long long threads = (long long)groups * 2 * p.rpw;
  hipLaunchKernelGGL(bat_best_argmin_kernel, dim3((unsigned)((threads + 63) / 64)), dim3(64), 0, st,
                     p, groups);
  return hipGetLastError();
}

}  // namespace blk

BatchedMfmaRun::~BatchedMfmaRun() {
  for (void* q : {frag, tab, pack, best_es})
    if (q) (void)hipFree(q);
}

hipError_t BatchedMfmaRun::enqueue(const BatchParams& bp, const std::vector<int>& inst,
                                   const int32_t* hn, const int32_t* hld, const int64_t* hj,
                                   const int64_t* hs, const int64_t* hh_off, const float* hscale,
                                   hipStream_t st) {
  using namespace blk;
  if (inst.empty()) return hipSuccess;
  const int R = bp.R, reads = bp.reads;
  if (R < 1 || R > 32) return hipErrorInvalidValue;
  const int rpt = 32 / R;  // reads per tile of 32 chains
  // Tiles (up to rpt reads of one instance), largest instances first; two
  // consecutive tiles of the same block count share a workgroup.
  int64_t foff = 0;
  int max_nblk = 0;
  pk.clear();
  wg.clear();
  std::vector<BatTile> tiles;
  std::vector<int> tnb;
  for (int k : inst) {
    const int n = hn[k], nblk = (n + kB - 1) / kB;
    if (n < kBatchedMfmaMinN || n > kBatchedMaxN) return hipErrorInvalidValue;
    BatPack b;
    b.n = n;
    b.nblk = nblk;
    b.ld = hld[k];
    b.pad = 0;
    b.sgn = hj[k];
    b.frag = foff;
    pk.push_back(b);
    max_nblk = nblk > max_nblk ? nblk : max_nblk;
    for (int r0 = 0; r0 < reads; r0 += rpt) {
      BatTile t;
      t.n = n;
      t.ld = hld[k];
      t.inst = k;
      t.read0 = k * reads + r0;
      t.read_end = k * reads + (r0 + rpt < reads ? r0 + rpt : reads);
      t.scale = hscale[k];
      t.frag = foff;
      t.sgn = hj[k];
      t.h = hh_off ? hh_off[k] : -1;
      t.spin_adj = hs[k] - (int64_t)k * reads * R * n;
      t.best_adj = hs[k] / R - (int64_t)k * reads * n;
      tiles.push_back(t);
      tnb.push_back(nblk);
    }
    foff += (int64_t)nblk * nblk * 64 * 16;
  }
  for (size_t i = 0; i < tiles.size();) {
    BatWG g;
    g.nblk = tnb[i];
    g.kb = (g.nblk + kPipeMW - 1) / kPipeMW;
    g.pad0 = g.pad1 = 0;
    g.t[0] = tiles[i];
    const bool pair = i + 1 < tiles.size() && tnb[i + 1] == tnb[i];
    g.t[1] = pair ? tiles[i + 1] : tiles[i];
    if (!pair) g.t[1].read_end = g.t[1].read0;  // an empty tile (valid pointers, no reads)
    wg.push_back(g);
    i += pair ? 2 : 1;
  }
  const int kbmax = (max_nblk + kPipeMW - 1) / kPipeMW;
  int nw = 4;
  while (2 * nw < kbmax) nw *= 2;
  const size_t chains = (size_t)bp.count * reads * R;
  hipError_t e = hipMalloc(&frag, (size_t)foff);
  if (e == hipSuccess) e = hipMalloc(&tab, sizeof(BatWG) * wg.size());
  if (e == hipSuccess) e = hipMalloc(&pack, sizeof(BatPack) * pk.size());
  if (e == hipSuccess && bp.best_sem == 0) e = hipMalloc(&best_es, chains * 8);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(tab, wg.data(), sizeof(BatWG) * wg.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pack, pk.data(), sizeof(BatPack) * pk.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = pack_batched_frags(bp.signs, (const BatPack*)pack, (int)pk.size(), max_nblk,
                           (int8_t*)frag, st);
  if (e != hipSuccess) return e;
  BParams p{};
  p.ptype = bp.ptype;
  p.R = R;
  p.reads = bp.count * reads;
  p.S = bp.S;
  p.rpw = rpt;  // reads per tile
  p.np = 1;
  p.use_pt = bp.use_pt;
  p.logw = 0;
  p.best_sem = bp.best_sem;
  p.order = bp.order;
  p.refresh = 0;  // integer counts: exact, nothing to recompute
  p.k0 = bp.k0;
  p.k1 = bp.k1;
  p.sweep_offset = bp.sweep_offset;
  p.frag = frag;
  p.h = bp.h;
  p.betas = bp.betas;
  p.spins = bp.states;
  p.energies = bp.energies;
  p.beta_idx = bp.beta_idx;
  p.best_rep = bp.best_rep;
  p.best_rep_e = (float*)best_es;
  p.best_rep_sw = best_es ? (int*)((char*)best_es + chains * 4) : nullptr;
  p.best_state = bp.best_state;
  p.best_energy = bp.best_energy;
  p.best_sweeps = bp.best_sweeps;
  p.log_omegas = nullptr;
  p.accepted = bp.accepted;
  p.bat = (const BatWG*)tab;
  p.sgn = bp.signs;
  e = launch_pipe_bat(p, nw, (int)wg.size(), st);
  if (e == hipSuccess && bp.best_sem == 0) e = bat_best_argmin(p, (int)wg.size(), st);
  return e;
}

}  // namespace pysa
