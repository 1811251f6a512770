// MI355X (gfx950) shallow-water step backend: kernels, device runtime, C ABI.
//
// One context owns the device copy of one mesh, its state, one HIP stream and
// the hipGraphs of recorded step blocks. A Heun step is five kernels in stream
// order (no host round trip):
//
//   k_stage<0>      predictor: per active 256-cell block, every edge touching
//                   the block is evaluated once into LDS, then each cell sums
//                   its edges in CSR order and updates (h1, hu1, hv1, h1_lo);
//   k_stage<1>      corrector: same on the predictor state, Heun average,
//                   the block's sources, wet mark, CFL min for the next step;
//   k_source_idle   sources of blocks the active set skips this step;
//   k_step_tail     in one launch, the dilation (active = wet over each
//                   block's two-ring, plus the list of blocks the next step
//                   visits) and the source fix (the exact dt bound of wetted
//                   source cells beside a dry cell of another block);
//   k_step_control  the uint64 tick clock and the next dt.
//
// The fused stage kernel replaces the reference's flux kernel + gather update
// (inundation/vulkan/shaders/resident.py:187-287 and :401-516): the per-edge
// words live in LDS instead of a 6-word-per-edge HBM buffer. An edge whose
// two sides are dry contributes zero words instead of being skipped; the sums
// start at +0.0, so adding +0.0 leaves them bit-identical to skipping.
//
// The C ABI (include/inundation_hip.h) returns 0 or a negative INU_E* code
// and keeps a per-context error string.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "swe_adjoint.h"
#include "swe_math.h"
#include "../../include/inundation_hip.h"

using inu::Phys;

namespace {

constexpr uint32_t kBlockCells = 256;  // ACTIVE_BLOCK_CELLS; one workgroup per block
constexpr uint32_t kXcds = 8;          // MI355X: 8 accelerator complex dies
constexpr uint32_t kWalkStride = 32;    // words between walk counters (one 128-byte line each)
constexpr uint32_t kGraphSteps = 100;   // steps per captured graph of inu_steps
constexpr float kTicksPerSecond = 1048576.0f;
constexpr float kDtSentinel = 1e30f;
constexpr uint32_t kNoBound = 0x7f800000u;  // +inf bits: larger than any bound
constexpr uint32_t kProbeOut = 7;           // outputs per inu_probe pair

// Clock and dt words of one scenario: the reference's 48-byte resident
// control block (gpu_resident_batch.py:65-77) and its dt_buf slots 0 and 1.
struct Scen {
  inu_ctrl ctrl;
  uint32_t dt_cur;   // this step's dt bits (0: no-op step); the prologue bound before the first
  uint32_t dt_next;  // next-step bound bits (atomicMin)
  uint32_t pad[2];
};
static_assert(sizeof(Scen) == 64, "scenario record must be 64 bytes");

// Control words shared by the step kernels and every scenario.
struct Ctl {
  uint32_t cnt_cur;    // blocks in `list` for the current step
  uint32_t cnt_next;   // blocks appended by k_dilate for the next step
  uint32_t swap_on;    // 1 when k_dilate runs every step (active set on)
  uint32_t idle_cur;   // inactive source blocks in `idle` for the current step
  uint32_t idle_next;  // inactive source blocks appended for the next step
  uint32_t any_dt;     // 1 when some scenario has a nonzero dt this step
  uint32_t pad[2];
  // Work of the active set, for accounting: {cells, owned edges, CSR slots}
  // of the active blocks of the current step, and summed over the real steps
  // taken (the next step's is in Args::act_part).
  unsigned long long act_cur[3];
  unsigned long long act_total[3];
  unsigned long long act_steps;
};
static_assert(sizeof(inu_ctrl) == 48, "control block must be 48 bytes");

// Device pointers and sizes, passed by value to every kernel.
// Where block b's data lives: its local edges (those with the left cell in b,
// then those with only the right cell in b) at [e0, e0 + n_loc) of the local
// edge stream, the first n_own of them owned; its ghost cells (the cells of
// other blocks that those edges touch) at gx[g0, g0 + n_ghost); its CSR slots
// at [s0, s0 + n_slots).
// meta = (scenario << 9) | number of real cells (1..256).
struct BlockInfo {
  uint32_t e0, n_loc, n_own, g0, n_ghost, s0, n_slots, meta;
};

__host__ __device__ __forceinline__ uint32_t bi_cells(const BlockInfo& bi) { return bi.meta & 511u; }
__host__ __device__ __forceinline__ uint32_t bi_scen(const BlockInfo& bi) { return bi.meta >> 9; }

struct Args {
  Phys p;
  uint32_t N, NB, max_local, max_slots, max_ghost, n_src, track_clamp;
  uint32_t n_lrec, n_lidx;  // entries of lrec / ldz and of lidx
  uint32_t src_neg;  // 1 when a source rate may be negative
  // cells
  float *h, *hu, *hv, *hlo, *h1, *hu1, *hv1, *h1lo, *area, *nman, *clamp, *src;
  float* inv_area;  // 1 / area, read by the fast order instead of area
  const uint32_t* cptr;
  // per cell: the end of its CSR slots within its block's slots (the stage
  // kernels' 16-bit copy of cptr[i + 1] - the block's first slot)
  const uint16_t* lend;
  const uint16_t* lidx;
  const int32_t* snb;
  const float* slen;
  // local edge streams of the blocks: {sides word, nx, ny, length} and the
  // signed bed step; ghost cell ids
  const uint4* lrec;
  const float* ldz;
  const BlockInfo* binfo;
  const uint32_t* gx;
  // blocks
  uint32_t *wet, *active, *forced, *has_src, *list, *idle;
  // 1 when a cell of the block may be below dry_tol after the last step's
  // sources, 0 when every cell is at or above it
  uint32_t* part_dry;
  const uint32_t *bptr, *bidx;
  // k_dilate's {cells, owned edges, CSR slots} of the next step's active
  // blocks, one triple per k_dilate workgroup (n_dwg of them)
  unsigned long long* act_part;
  uint32_t n_dwg;
  // sources
  const uint32_t* src_cell;
  const uint32_t* src_blocks;
  // (source cell, its neighbour in another block), grouped by the
  // neighbour's block: the neighbours of group f lie in block fix_blk[f];
  // the pairs [kFixChunk w, kFixChunk (w + 1)) touch the groups
  // [fix_chunk[w].x, fix_chunk[w].y)
  const uint2* src_pair;
  const uint32_t* fix_blk;
  const uint2* fix_chunk;
  uint32_t n_src_blocks, n_src_pairs, n_fix;
  Ctl* ctl;
  Scen* scen;
  // The walk counters of the stage kernels and of the adjoint's reverse
  // stages: per kernel, kXcds slice counters, kWalkStride words apart; zero
  // between launches.
  uint32_t* walk;
  uint32_t n_scen;
  // open boundary edges (device record, read only where they are handled)
  const struct OpenDev* open;
  uint32_t n_open;
  // adjoint tape (recording stage kernels): kTapeWords arrays of N per step
  // slot; the slot is the replay cursor's step minus its segment start
  float* tape;
  const struct AdjCur* adj_cur;
};

// Open boundary edges: condition and value per open edge; each open edge's
// six words per stage (from k_open_words) and its mass word (flux x length)
// per stage, for the boundary tally; per block, its open edges
// [bptr[b], bptr[b + 1]) as (local edge position, open edge index) (the edge
// stream holds them as walls).
struct OpenDev {
  const uint32_t* kind;
  const float* value;
  float4* words;
  float* flux;
  const uint32_t* bptr;
  const uint2* jk;
};

// Device-side cursor of an adjoint replay or reverse pass: the step being
// replayed or reversed, the first step of its segment (tape slot = k - k0),
// and counters of the reverse pass: listed block-steps, idle source
// block-steps, and edges whose gradient reached a cell outside the listed
// blocks (film on the active set's boundary).
struct AdjCur {
  uint32_t k, k0;
  unsigned long long stats[3];
};

// Tape words per cell and recorded step: stage 0 input (h, hu, hv), its
// updated state before friction (h', hu', hv') and clamp flag; the same for
// stage 1.
constexpr uint32_t kTapeWords = 14;

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = (o < v) ? o : v;
  }
  return v;
}

// Min of a per-thread float bound (positive or kNoBound) over the workgroup
// (blockDim.x <= 256), then one atomicMin: positive floats order like their
// bit patterns. Every thread of the workgroup must call it.
__device__ __forceinline__ void block_atomic_min(uint32_t* dst, uint32_t bits) {
  __shared__ uint32_t wave_min[4];
  const uint32_t m = wave_min_u32(bits);
  const uint32_t w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63u) == 0u) wave_min[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t v = wave_min[0];
    for (uint32_t k = 1; k < (blockDim.x + 63u) / 64u; ++k) v = (wave_min[k] < v) ? wave_min[k] : v;
    // Most workgroups do not hold the minimum: read before the atomic.
    if (v != kNoBound && v < __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMin(dst, v);
  }
}

// ---------------------------------------------------------------------------
// Stage kernels
// ---------------------------------------------------------------------------

// Element idx of a device array addressed as base + a 32-bit byte offset
// (every array is below 4 GiB), so the address is a scalar base plus one
// vector offset shared by the arrays read at the same index.
// dt bound bits of a cell over its summed wave speeds lam: cfl A / lam for
// lam > 1e-30 and A > 1e-10 (reference resident.py:509, :604, :726), else
// none. The fast order reads the cell's inverse area aw = 1 / A and takes
// cfl / (aw lam), guarded by aw < 1e10; the exact order reads aw = A.
template <bool FM>
__device__ __forceinline__ uint32_t cell_bound(const Phys& p, float aw, float lam) {
  if (!(lam > 1e-30f)) return kNoBound;
  if (FM) return aw < 1e10f ? __float_as_uint(p.cfl / (aw * lam)) : kNoBound;
  return aw > 1e-10f ? __float_as_uint(p.cfl * aw / lam) : kNoBound;
}

template <typename T>
__device__ __forceinline__ T ldg(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (idx * (uint32_t)sizeof(T)));
}
// A load through the constant address space, which the compiler issues as a
// scalar load when the index is uniform: no vector registers and no wait on
// the vector memory counter. Only for data the running kernel never writes
// (the scalar cache is invalidated at every kernel launch).
__device__ __forceinline__ uint32_t ldc(const uint32_t* base, uint32_t idx) {
  typedef const __attribute__((address_space(4))) uint32_t* cptr_t;
  return ((cptr_t)base)[idx];
}
__device__ __forceinline__ BlockInfo ldc(const BlockInfo* base, uint32_t idx) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + 8 * (size_t)idx;
  return BlockInfo{ldc(w, 0), ldc(w, 1), ldc(w, 2), ldc(w, 3), ldc(w, 4), ldc(w, 5), ldc(w, 6), ldc(w, 7)};
}
template <typename T>
__device__ __forceinline__ void stg(T* base, uint32_t idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (idx * (uint32_t)sizeof(T))) = v;
}

// LDS of one block. Each cell and ghost cell has a record of its stage input
// (h, hu, hv, lo) in the exact order, or (h, u, v, lo) with the
// desingularised velocity and, beside it, (phi(h), sqrt h) in the fast order.
// Each local edge has one 32-byte record: its words {w0, w1, w2, w3}, {w4, w5,
// sides word, length}. Then the block's CSR slots (local edge numbers).
struct Lds {
  float4* cell;  // the cells, then the ghosts: [kBlockCells + max_ghost][2] in
                 // carve() (prologue); the stage kernels' first record only
  float2* cx;    // the stage kernels' fast-order (phi(h), sqrt h) per cell
  float4* edge;  // [max_local][2] (prologue, adjoint); the corrector's side records
  float2* e2;    // carve(): words 4-5 beside edge
  float* ps;     // the predictor's side records, 3 words per CSR slot value
  float* elen;   // the corrector's edge lengths, for cell perimeters
  uint16_t* s_slot;
};

// Local cell numbers of an edge's sides word ((left << 16) | right): a cell of
// the block (0..255), a ghost (256 + its place in the block's ghost list) or,
// on the right only, a wall.
constexpr uint32_t kSideWall = 0xffffu;
constexpr uint32_t kMaxGhosts = kSideWall - kBlockCells;

// Tile levels of the stage kernels, chosen per mesh by inu_autotune: edge
// records and CSR slots each thread loads before the first barrier (the rest
// are loaded after it). Denser meshes (Voronoi cells have about three edges
// per cell) need more to keep a block to one memory round trip; more costs
// registers.
constexpr int kTiles = 3;
__host__ __device__ constexpr uint32_t edges_ahead(int tile) { return tile == 0 ? 2u : tile == 1 ? 3u : 4u; }
__host__ __device__ constexpr uint32_t slots_ahead(int tile) { return tile == 0 ? 4u : tile == 1 ? 6u : 8u; }

__host__ __device__ inline size_t lds_words(uint32_t max_local, uint32_t max_slots, uint32_t max_ghost) {
  return 8 * (size_t)(kBlockCells + max_ghost) + 8 * (size_t)max_local + (max_slots + 1) / 2;
}

__device__ __forceinline__ Lds carve(float* sm, uint32_t max_local, uint32_t max_ghost) {
  Lds l;
  l.cell = reinterpret_cast<float4*>(sm);
  l.edge = reinterpret_cast<float4*>(sm + 8 * (kBlockCells + max_ghost));
  l.e2 = reinterpret_cast<float2*>(l.edge) + 2;
  l.s_slot = reinterpret_cast<uint16_t*>(sm + 8 * (kBlockCells + max_ghost) + 8 * max_local);
  return l;
}

// The stage kernels keep 24 bytes per cell: (h, u, v, lo) and, for the fast
// order, (phi(h), sqrt h) in a second array; the pressure half_g h^2 is
// formed again where an edge needs it. Cell count rounded up to even, so the
// edge records after them stay 16-byte aligned.
__host__ __device__ inline uint32_t stage_cells(uint32_t max_ghost) { return (kBlockCells + max_ghost + 1u) & ~1u; }

// The stage kernels keep an edge's words as one record per side, indexed by
// the CSR slot value s = (j << 1) | right of local edge j, so each slot of
// the gather reads its own contribution: mass, momentum x, momentum y as
// (w0, w1, w2) on the left and (-w0, -w3, -w4) on the right. The predictor
// keeps these 3 words per side (24 bytes per edge); the corrector adds the
// wave speed term w5 to each (two 16-byte records, 32 bytes per edge) and
// keeps each edge's length beside them, from which the all-wet source bound
// sums a cell's perimeter.
__host__ __device__ inline size_t lds_words_stage0(uint32_t max_local, uint32_t max_slots, uint32_t max_ghost) {
  return 6 * (size_t)stage_cells(max_ghost) + 6 * (size_t)max_local + (max_slots + 1) / 2;
}
__host__ __device__ inline size_t lds_words_stage1(uint32_t max_local, uint32_t max_slots, uint32_t max_ghost) {
  return 6 * (size_t)stage_cells(max_ghost) + 9 * (size_t)max_local + (max_slots + 1) / 2;
}

template <int ST>
__device__ __forceinline__ Lds carve_stage(float* sm, uint32_t max_local, uint32_t max_ghost) {
  const uint32_t kc = stage_cells(max_ghost);
  Lds l;
  l.cell = reinterpret_cast<float4*>(sm);
  l.cx = reinterpret_cast<float2*>(sm + 4 * kc);
  float* e = sm + 6 * kc;
  l.edge = reinterpret_cast<float4*>(e);
  l.e2 = nullptr;
  l.ps = e;
  l.elen = e + 8 * max_local;  // the corrector's only
  l.s_slot = reinterpret_cast<uint16_t*>(e + (ST == 0 ? 6 : 9) * max_local);
  return l;
}

// The side records of local edge j from its six words w.
template <int ST>
__device__ __forceinline__ void put_sides(const Lds& l, uint32_t j, const float w[6]) {
  if (ST == 0) {
    float2* q = reinterpret_cast<float2*>(l.ps + 6 * j);
    q[0] = make_float2(w[0], w[1]);
    q[1] = make_float2(w[2], -w[0]);
    q[2] = make_float2(-w[3], -w[4]);
  } else {
    l.edge[2 * j] = make_float4(w[0], w[1], w[2], w[5]);
    l.edge[2 * j + 1] = make_float4(-w[0], -w[3], -w[4], w[5]);
  }
}

// One local edge record: sides word, normal, length and the signed bed step
// (dzL when positive, -dzR when negative: one of the two is always zero).
struct EdgeRec {
  uint32_t side;
  float nx, ny, len, d;
};

// The record at position q of the local edge streams.
__device__ __forceinline__ EdgeRec load_edge_at(const Args& a, uint32_t q) {
  const uint4 v = ldg(a.lrec, q);
  EdgeRec r;
  r.side = v.x;
  r.nx = __uint_as_float(v.y);
  r.ny = __uint_as_float(v.z);
  r.len = __uint_as_float(v.w);
  r.d = ldg(a.ldz, q);
  return r;
}

__device__ __forceinline__ EdgeRec load_edge(const Args& a, const BlockInfo& bi, uint32_t j) {
  return load_edge_at(a, bi.e0 + j);
}

__device__ __forceinline__ uint32_t umin(uint32_t x, uint32_t y) { return x < y ? x : y; }

// Stage input of ghost cell g (a cell of another block). The corrector reads
// the predictor state only where that cell's block was active this step
// (reference resident.py:209-226); it loads the block's flag and both states
// at once and keeps one, so no load waits for another.
template <int ST, bool LO>
__device__ __forceinline__ float4 load_ghost(const Args& a, uint32_t g) {
  const float4 s0 = make_float4(ldg(a.h, g), ldg(a.hu, g), ldg(a.hv, g), LO ? ldg(a.hlo, g) : 0.0f);
  if (ST == 0) return s0;
  const uint32_t act = ldg(a.active, g >> 8);
  const float4 s1 = make_float4(ldg(a.h1, g), ldg(a.hu1, g), ldg(a.hv1, g), LO ? ldg(a.h1lo, g) : 0.0f);
  return (act != 0u) ? s1 : s0;
}

// A cell's step-start (h, hu, hv, lo) and predictor-stage state.
template <bool LO>
__device__ __forceinline__ float4 state0(const Args& a, uint32_t g) {
  return make_float4(ldg(a.h, g), ldg(a.hu, g), ldg(a.hv, g), LO ? ldg(a.hlo, g) : 0.0f);
}
template <bool LO>
__device__ __forceinline__ float4 state1(const Args& a, uint32_t g) {
  return make_float4(ldg(a.h1, g), ldg(a.hu1, g), ldg(a.hv1, g), LO ? ldg(a.h1lo, g) : 0.0f);
}

// Adjoint tape: a ghost cell whose block is not listed this step keeps its
// step state through both stages; its stage input is recorded for the reverse
// pass (listed cells are recorded by their own block).
template <int ST, bool COMP>
__device__ __forceinline__ void rec_ghost(const Args& a, uint32_t g, float4 s) {
  if (ldg(a.active, g >> 8) != 0u) return;
  const uint32_t slot = a.adj_cur->k - a.adj_cur->k0;
  float* t = a.tape + (size_t)(slot * kTapeWords + (ST == 0 ? 0u : 7u)) * a.N + g;
  t[0] = COMP ? ldg(a.h, g) + ldg(a.hlo, g) : s.x;
  t[a.N] = s.y;
  t[2 * (size_t)a.N] = s.z;
}

// Stage input s = (h, hu, hv, lo) of one cell or ghost into LDS record q. The
// exact order keeps it as is; the fast order keeps (h, u, v, lo) with the
// desingularised velocity, and (phi(h), sqrt h) beside it.
template <bool FM>
__device__ __forceinline__ void put_cell(const Lds& l, const Phys& p, uint32_t q, float4 s) {
  if (FM) {
    const inu::fast::Cell c = inu::fast::cell(p, s.x, s.y, s.z, s.w);
    l.cell[q] = make_float4(c.h, c.u, c.v, c.lo);
    l.cx[q] = make_float2(c.phi, c.sq);
  } else {
    l.cell[q] = s;
  }
}

// The fast-order side of LDS record q.
__device__ __forceinline__ inu::fast::Cell get_cell(const Lds& l, uint32_t q) {
  const float4 s = l.cell[q];
  const float2 x = l.cx[q];
  return inu::fast::Cell{s.x, s.y, s.z, s.w, x.x, x.y};
}

// The words of an open edge from its inside cell sl = (h, hu, hv) with the
// state (hR, huR, hvR) beyond it; `wall` passes no mass; fh receives the HLLC
// mass flux. The face depths are the cell depths (no bed step).
template <bool FM>
__device__ __forceinline__ void open_words(const Phys& p, float4 sl, float hR, float huR, float hvR, bool wall,
                                           const EdgeRec& r, float w[6], float& fh) {
  if (FM) {
    const inu::fast::Cell L = inu::fast::cell(p, sl.x, sl.y, sl.z, 0.0f);
    const inu::fast::Cell R = inu::fast::cell(p, hR, huR, hvR, 0.0f);
    inu::fast::edge_words(p, wall, L, R, false, 0.0f, r.nx, r.ny, r.len, w, &fh);
  } else {
    inu::edge_words(p, wall, sl.x, sl.y, sl.z, hR, huR, hvR, 0.0f, 0.0f, r.nx, r.ny, r.len, w, &fh);
  }
}

// Words of an open edge (reference resident.py:130-142 and flux.py:271-296)
// from its inside cell's stage input sl = (h, hu, hv): the ghost state of its
// condition; a dry
// edge passes nothing unless an inflow keeps it live; an outflow-only edge
// whose flux would draw water in is redone against the wall ghost; otherwise
// the mass word is the condition's boundary flux.
template <bool FM>
__device__ void open_edge_words(const Phys& p, uint32_t kind, float value, float4 sl, const EdgeRec& r, float w[6]) {
  float hG, huG, hvG;
  inu::ghost_state(p, kind, value, sl.x, sl.y, sl.z, r.nx, r.ny, hG, huG, hvG);
  for (int q = 0; q < 6; ++q) w[q] = 0.0f;
  if (!(sl.x < p.dry && hG < p.dry) || inu::open_forced(kind, value)) {
    float fh;
    open_words<FM>(p, sl, hG, huG, hvG, false, r, w, fh);
    if (inu::open_outflow_only(kind) && fh < 0.0f) {
      inu::ghost_state(p, inu::kOpenWall, 0.0f, sl.x, sl.y, sl.z, r.nx, r.ny, hG, huG, hvG);
      open_words<FM>(p, sl, hG, huG, hvG, true, r, w, fh);
    } else {
      w[0] = inu::boundary_mass(kind, value, fh) * r.len;
    }
  }
}

// Open edge words of stage ST, one thread per open edge, before the stage
// kernel (which copies them into its LDS edge list), from the inside cell's
// stage input as the stage kernel loads it; the mass word also goes to the
// boundary tally. A cell whose block is not visited gets words nobody reads.
template <int ST, bool FM>
__global__ void k_open_words(const Args a, const uint32_t* open_cell, const float4* open_geom) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n_open || a.ctl->any_dt == 0u) return;
  const Phys& p = a.p;
  const uint32_t i = open_cell[k];
  const float h = (ST == 0) ? a.h[i] : a.h1[i];
  const float4 sl = make_float4(h, (ST == 0) ? a.hu[i] : a.hu1[i], (ST == 0) ? a.hv[i] : a.hv1[i], 0.0f);
  const float4 g = open_geom[k];
  EdgeRec r;
  r.side = 0u;
  r.nx = g.x;
  r.ny = g.y;
  r.len = g.z;
  r.d = 0.0f;
  float w[6];
  open_edge_words<FM>(p, a.open->kind[k], a.open->value[k], sl, r, w);
  float4* o = a.open->words + 2 * ((size_t)ST * a.n_open + k);
  o[0] = make_float4(w[0], w[1], w[2], w[3]);
  o[1] = make_float4(w[4], w[5], 0.0f, 0.0f);
  a.open->flux[(size_t)ST * a.n_open + k] = w[0];
}

// Open edges of block b, after its edge list is complete: their words from
// k_open_words replace the wall words. Only a mesh with open edges pays for
// the pass.
template <int ST>
__device__ __forceinline__ void open_pass(const Args& a, const Lds& l, uint32_t b) {
  const uint32_t t1 = a.open->bptr[b + 1];
  for (uint32_t t = a.open->bptr[b] + threadIdx.x; t < t1; t += kBlockCells) {
    const uint2 jk = a.open->jk[t];
    const float4* o = a.open->words + 2 * ((size_t)ST * a.n_open + jk.y);
    const float4 w0 = o[0], w1 = o[1];
    const float w[6] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y};
    put_sides<ST>(l, jk.x, w);
  }
}

// The six words of a local edge (zeros when both sides are dry), both sides
// read from LDS. An open edge is stored as a wall; open_pass replaces its
// words.
template <bool FM, bool LO>
__device__ __forceinline__ void edge_words_lds(const Args& a, const Lds& l, const EdgeRec& r, float w[6]) {
  const uint32_t ql = r.side >> 16;
  const uint32_t qr = r.side & 0xffffu;
  const bool wall = qr == kSideWall;
  for (int q = 0; q < 6; ++q) w[q] = 0.0f;
  if (FM) {
    // A wall's right side is the left cell's mirror.
    const inu::fast::Cell L = get_cell(l, ql);
    const inu::fast::Cell R = wall ? L : get_cell(l, qr);
    if (!(L.h < a.p.dry && R.h < a.p.dry)) {
      const inu::fast::Cell Lo{L.h, L.u, L.v, LO ? L.lo : 0.0f, L.phi, L.sq};
      const inu::fast::Cell Ro{R.h, R.u, R.v, LO ? R.lo : 0.0f, R.phi, R.sq};
      inu::fast::edge_words(a.p, wall, Lo, Ro, wall, r.d, r.nx, r.ny, r.len, w);
    }
    return;
  }
  const float dzL = (r.d > 0.0f) ? r.d : 0.0f;
  const float dzR = (r.d < 0.0f) ? -r.d : 0.0f;
  const float4 sl = l.cell[ql];
  float4 sr;
  if (wall) {
    sr = sl;
    inu::wall_ghost(sl.y, sl.z, r.nx, r.ny, sr.y, sr.z);
  } else {
    sr = l.cell[qr];
  }
  if (!(sl.x < a.p.dry && sr.x < a.p.dry))
    inu::edge_words(a.p, wall, sl.x, sl.y, sl.z, sr.x, sr.y, sr.z, dzL, dzR, r.nx, r.ny, r.len, w);
}

// Speed |u| and celerity c of a wetted source cell of depth h for its dt bound
// (reference resident.py:709-728).
template <bool FM>
__device__ __forceinline__ void src_uc(const Phys& p, float h, float hu, float hv, float& u, float& c) {
  if (FM) {
    inu::fast::src_speeds(p, h, hu, hv, u, c);
  } else {
    u = inu::sqrt_rn(hu * hu + hv * hv) / h;
    c = inu::sqrt_rn(p.g * h);
  }
}

// The side records of local edge j: left lw and right rw (mass, momentum x,
// momentum y, wave speed term), the right side's first three negated.
template <int ST>
__device__ __forceinline__ void put_sides4(const Lds& l, uint32_t j, float4 lw, float4 rw) {
  if (ST == 0) {
    float2* q = reinterpret_cast<float2*>(l.ps + 6 * j);
    q[0] = make_float2(lw.x, lw.y);
    q[1] = make_float2(lw.z, rw.x);
    q[2] = make_float2(rw.y, rw.z);
  } else {
    l.edge[2 * j] = lw;
    l.edge[2 * j + 1] = rw;
  }
}

// Words of local edge j into LDS, for the fixed-order gather. In the fast
// order a wall reads its left cell twice (its right side is that cell's
// mirror), and an edge dry on both sides stores zero words (the negated
// right words as -0, as negating zeros would give).
template <int ST, bool FM, bool LO>
__device__ __forceinline__ void edge_eval(const Args& a, const Lds& l, const EdgeRec& r, uint32_t j) {
  if (ST == 1) l.elen[j] = r.len;
  if (!FM) {
    float w[6];
    edge_words_lds<FM, LO>(a, l, r, w);
    put_sides<ST>(l, j, w);
    return;
  }
  const uint32_t ql = r.side >> 16;
  const bool wall = (r.side & 0xffffu) == kSideWall;
  float4 sl = l.cell[ql], sr = l.cell[wall ? ql : (r.side & 0xffffu)];
  float2 xl = l.cx[ql], xr = l.cx[wall ? ql : (r.side & 0xffffu)];
  // Both records of each side in whole reads, before the dry test.
  asm volatile("" : "+v"(sl.w), "+v"(sr.w), "+v"(xl.x), "+v"(xl.y), "+v"(xr.x), "+v"(xr.y));
  const inu::fast::Cell L{sl.x, sl.y, sl.z, sl.w, xl.x, xl.y};
  const inu::fast::Cell R{sr.x, sr.y, sr.z, sr.w, xr.x, xr.y};
  if (L.h < a.p.dry && R.h < a.p.dry) {
    put_sides4<ST>(l, j, make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(-0.0f, -0.0f, -0.0f, 0.0f));
    return;
  }
  float4 lw, rw;
  inu::fast::edge_sides(a.p, wall, L, R, wall, r.d, r.nx, r.ny, r.len, lw, rw);
  put_sides4<ST>(l, j, lw, rw);
}

// Unordered accumulation (the reference's atomic flux variant,
// resident.py:110-180): the edge's words are added at once into the LDS sums
// of its sides in this block (mass, momentum x, y, wave speeds), in whatever
// order the threads arrive, so results vary in the last bits from run to run.
// Ghost sides are summed by their own block.
template <int ST, bool FM, bool LO>
__device__ __forceinline__ void edge_accumulate(const Args& a, const Lds& l, const EdgeRec& r) {
  float w[6];
  edge_words_lds<FM, LO>(a, l, r, w);
  if (w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f && w[4] == 0.0f && w[5] == 0.0f) return;
  // Four sums per cell, each a run of 256 words (one LDS bank per cell).
  float* acc = reinterpret_cast<float*>(l.edge);
  const uint32_t ql = r.side >> 16;
  const uint32_t qr = r.side & 0xffffu;
  if (ql < kBlockCells) {
    atomicAdd(acc + ql, w[0]);
    atomicAdd(acc + kBlockCells + ql, w[1]);
    atomicAdd(acc + 2 * kBlockCells + ql, w[2]);
    if (ST == 1) atomicAdd(acc + 3 * kBlockCells + ql, w[5]);
  }
  if (qr < kBlockCells) {
    atomicAdd(acc + qr, -w[0]);
    atomicAdd(acc + kBlockCells + qr, -w[3]);
    atomicAdd(acc + 2 * kBlockCells + qr, -w[4]);
    if (ST == 1) atomicAdd(acc + 3 * kBlockCells + qr, w[5]);
  }
}

// The next blocks of a workgroup's walk over the list: the block after the
// current one with its BlockInfo and this thread's ghost cell id, and the
// list entry after that. Loaded while the current block is processed.
struct Ahead {
  uint32_t t;    // list position of the next block (>= end: none)
  uint32_t b;    // next block
  BlockInfo bi;  // its BlockInfo
  uint32_t gid;  // its ghost cell id of this thread
  uint32_t dtw;  // its scenario's dt bits
  uint32_t t2;   // list position of the block after it (>= end: none)
  uint32_t b2;   // block after it
};

constexpr uint32_t kNoPos = 0xffffffffu;  // no list position

// A workgroup's share of the walk (three words, kept in scalar registers).
// Its slice's list positions from `base` on, `len` of them, are handed out
// by the slice's counter (those before, two per workgroup, are the
// workgroups' first blocks; a position past the slice is kNoPos). A
// workgroup takes positions until one take comes back empty, and never
// takes again (`last` becomes kNoPos), so the counter ends at exactly the
// initial `last` + 1: the handed out positions plus one empty take per
// workgroup that took any. The take that returns that `last` is the
// slice's final one and clears the counter for the next launch.
struct Walk {
  uint32_t base, len, last;
};

// The counter of this workgroup's slice in stage ST's launch.
template <int ST>
__device__ __forceinline__ uint32_t* walk_ctr(const Args& a) {
  return a.walk + ((uint32_t)ST * kXcds + blockIdx.x % kXcds) * kWalkStride;
}

// The stage kernels' LDS word that passes a taken position from the thread
// that took it to the workgroup.
__shared__ uint32_t s_walk_pos;

// Position of a take that returned counter value r (kNoPos: an empty take,
// or none); thread 0 clears the counter after the slice's final take.
template <int ST>
__device__ __forceinline__ uint32_t walk_pos(const Args& a, const Walk& wk, uint32_t r) {
  if (threadIdx.x == 0u && r != kNoPos && r == wk.last) atomicExch(walk_ctr<ST>(a), 0u);
  return (r < wk.len) ? wk.base + r : kNoPos;
}

// One active block of one Heun stage (reference resident.py:425-512, gather +
// compensated variants), with the corrector also applying this block's
// sources (resident.py:691-708). Three phases: (A) every load of the block is
// issued at once (cells, ghost cells, CSR slots, the thread's first edge
// records); (B) each edge touching the block is evaluated once into LDS from
// LDS; (C) each cell sums its edges in CSR order from LDS and updates.
template <int ST, bool COMP, bool FM, int TILE, bool REC, bool ATOM, bool XTRA>
__device__ uint32_t stage_block(const Args& a, uint32_t b, const BlockInfo& bi, uint32_t gid, float dt,
                                const Lds& l, Ahead& nx, const Walk& wk) {
  constexpr bool LO = COMP && FM;  // the fast order's face depths read the low words
  constexpr uint32_t kEdgesAhead = edges_ahead(TILE);
  constexpr uint32_t kSlotsAhead = slots_ahead(TILE);
  const Phys& p = a.p;
  const uint32_t tid = threadIdx.x;
  const uint32_t c0 = b << 8;
  const uint32_t nc = bi_cells(bi);
  const uint32_t n_loc = bi.n_loc;

  // ---- (A) independent loads -------------------------------------------
  // One batch of loads, none waiting for another: a group of loads is
  // skipped only by waves with no lane in its range (a branch uniform over
  // the wave), and lanes past the block's edges, slots or ghosts read a
  // valid neighbouring entry whose value is not used.
  const uint32_t wave0 = __builtin_amdgcn_readfirstlane(tid & ~63u);  // the wave's first thread
  // The ghost this thread loads: taken first, since its id came from the
  // last block's loads (so no load of this batch waits behind it).
  uint32_t gsrc = (tid < bi.n_ghost) ? gid : c0;
  asm volatile("" : "+v"(gsrc));
  // The corrector reads a ghost's predictor state only where its block was
  // active this step (reference resident.py:209-226): the flag is the
  // batch's first load, so the rare start-state load it calls for, issued
  // last, waits on it alone.
  const uint32_t gact = (ST == 1) ? ldg(a.active, gsrc >> 8) : 1u;
  // Walk ahead (scalar loads): the next block's BlockInfo.
  if (nx.t != kNoPos) nx.bi = ldc(a.binfo, nx.b);
  EdgeRec er[kEdgesAhead];
#pragma unroll
  for (uint32_t m = 0; m < kEdgesAhead; ++m)
    if (wave0 + m * kBlockCells < n_loc) er[m] = load_edge_at(a, umin(bi.e0 + tid + m * kBlockCells, a.n_lrec - 1u));
  uint32_t sl[kSlotsAhead];
#pragma unroll
  for (uint32_t m = 0; m < kSlotsAhead; ++m)
    if (!ATOM && wave0 + m * kBlockCells < bi.n_slots) sl[m] = ldg(a.lidx, umin(bi.s0 + tid + m * kBlockCells, a.n_lidx - 1u));
  // The ghost's stage input: the predictor state in the corrector (its
  // block's flag, loaded first, picks the start state instead below).
  float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (wave0 < bi.n_ghost) gs = (ST == 1) ? state1<LO>(a, gsrc) : state0<LO>(a, gsrc);
  const uint32_t i = c0 + tid;  // the arrays are padded to whole blocks
  // aw: the cell's area, or its inverse in the fast order
  const float hi0 = ldg((ST == 0) ? a.h : a.h1, i);
  const float hui0 = ldg((ST == 0) ? a.hu : a.hu1, i);
  const float hvi0 = ldg((ST == 0) ? a.hv : a.hv1, i);
  const float loi0 = COMP ? ldg((ST == 0) ? a.hlo : a.h1lo, i) : 0.0f;
  // The cell's slots [k0, k1) within the block's: 16-bit ends, the first
  // cell's start is 0.
  const uint32_t k1 = ATOM ? 0u : (uint32_t)ldg(a.lend, i);
  const uint32_t k0 = (ATOM || tid == 0) ? 0u : (uint32_t)ldg(a.lend, i - 1u);
  const float aw = ldg(FM ? a.inv_area : a.area, i);
  const float nm = ldg(a.nman, i);
  float h0 = 0.0f, hu0 = 0.0f, hv0 = 0.0f, lo0 = 0.0f, rate = 0.0f;
  if (ST == 1) {
    h0 = ldg(a.h, i);
    hu0 = ldg(a.hu, i);
    hv0 = ldg(a.hv, i);
    if (COMP) lo0 = ldg(a.hlo, i);
    if (a.n_src != 0u) rate = ldg(a.src, i);
  }
  // Walk ahead, last in the batch so nothing below waits for it: this
  // thread's ghost cell id of the next block, and its scenario's dt.
  if (nx.t != kNoPos) {
    if (tid < nx.bi.n_ghost) nx.gid = ldg(a.gx, nx.bi.g0 + tid);
    nx.dtw = ldc(&a.scen[bi_scen(nx.bi)].dt_cur, 0u);
  }
  if (ST == 1 && tid < bi.n_ghost && gact == 0u) gs = state0<LO>(a, gsrc);
  // The list position of the block after the next one, taken from the
  // slice's counter after every load above (nothing before phase C waits
  // for it). An increment rather than an add: the compiler's wave-wide
  // rewrite of a uniform add would wait for its return at once.
  uint32_t taken = kNoPos;
  if (tid == 0u && wk.last != kNoPos) taken = atomicInc(walk_ctr<ST>(a), 0xffffffffu);
  float hi = hi0, hui = hui0, hvi = hvi0;
  const float loi = loi0;
  __syncthreads();  // the previous block's LDS reads are done
  if (tid < nc) put_cell<FM>(l, p, tid, make_float4(hi, hui, hvi, loi));
  if (tid < bi.n_ghost) {
    put_cell<FM>(l, p, kBlockCells + tid, gs);
    if (REC) rec_ghost<ST, COMP>(a, gid, gs);
  }
  for (uint32_t g = tid + kBlockCells; g < bi.n_ghost; g += kBlockCells) {
    const uint32_t gg = ldg(a.gx, bi.g0 + g);
    const float4 s = load_ghost<ST, LO>(a, gg);
    put_cell<FM>(l, p, kBlockCells + g, s);
    if (REC) rec_ghost<ST, COMP>(a, gg, s);
  }
  if (ATOM) {
    float* acc = reinterpret_cast<float*>(l.edge);
    for (uint32_t k = 0; k < 4; ++k) acc[k * kBlockCells + tid] = 0.0f;
  } else {
#pragma unroll
    for (uint32_t m = 0; m < kSlotsAhead; ++m)
      if (tid + m * kBlockCells < bi.n_slots) l.s_slot[tid + m * kBlockCells] = (uint16_t)sl[m];
    for (uint32_t k = tid + kSlotsAhead * kBlockCells; k < bi.n_slots; k += kBlockCells)
      l.s_slot[k] = ldg(a.lidx, bi.s0 + k);
  }
  __syncthreads();

  // ---- (B) edges -----------------------------------------------------------
#pragma unroll
  for (uint32_t m = 0; m < kEdgesAhead; ++m) {
    if (tid + m * kBlockCells < n_loc) {
      if (ATOM) edge_accumulate<ST, FM, LO>(a, l, er[m]);
      else edge_eval<ST, FM, LO>(a, l, er[m], tid + m * kBlockCells);
    }
  }
  for (uint32_t j = tid + kEdgesAhead * kBlockCells; j < n_loc; j += kBlockCells) {
    if (ATOM) edge_accumulate<ST, FM, LO>(a, l, load_edge(a, bi, j));
    else edge_eval<ST, FM, LO>(a, l, load_edge(a, bi, j), j);
  }
  if (tid == 0u) s_walk_pos = taken;
  __syncthreads();
  if (XTRA && a.n_open != 0u) {
    open_pass<ST>(a, l, b);
    __syncthreads();
  }

  // ---- (C) cells -----------------------------------------------------------
  // The list entry at the position taken in phase A: a vector load, so the
  // LDS reads below do not wait on it.
  nx.t2 = walk_pos<ST>(a, wk, __builtin_amdgcn_readfirstlane(s_walk_pos));
  if (nx.t2 != kNoPos) nx.b2 = ldg(a.list, nx.t2);
  bool wet = false;
  uint32_t bound = kNoBound;
  // The corrector's dt bound is cfl A / lam over the cell's wave speeds and,
  // for a wetted source cell, over its source speeds: the smaller bound is
  // the one over the larger lam (the same rounded numerator), so one
  // division takes both.
  float lam_b = 0.0f;
  // The corrector's fast order takes the source dt bound of all-wet blocks
  // from the cells' perimeters.
  constexpr bool PERIM = ST == 1 && FM && !ATOM;
  float perim = 0.0f;  // PERIM with sources: the cell's edge lengths summed in CSR order
  if (tid < nc) {
    float sdh = 0.0f, sdhu = 0.0f, sdhv = 0.0f, lam = 0.0f;
    if (ATOM) {
      const float* acc = reinterpret_cast<const float*>(l.edge);
      sdh = acc[tid];
      sdhu = acc[kBlockCells + tid];
      sdhv = acc[2 * kBlockCells + tid];
      lam = acc[3 * kBlockCells + tid];
    }
    // Each slot adds its side's record: a right side's words are stored
    // negated, and x + (-y) is x - y bit for bit.
    if (PERIM && a.n_src != 0u) {
      for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t s = l.s_slot[k];
        const float4 w = l.edge[s];
        sdh += w.x;
        sdhu += w.y;
        sdhv += w.z;
        lam += w.w;
        perim += l.elen[s >> 1];
      }
    }
    for (uint32_t k = k0; !ATOM && !(PERIM && a.n_src != 0u) && k < k1; ++k) {
      const uint32_t s = l.s_slot[k];
      if (ST == 1) {
        const float4 w = l.edge[s];
        sdh += w.x;
        sdhu += w.y;
        sdhv += w.z;
        lam += w.w;
      } else {
        const float* q = l.ps + 3 * s;
        sdh += q[0];
        sdhu += q[1];
        sdhv += q[2];
      }
    }
    const float dti = FM ? dt * aw : dt / aw;
    float h_new, lo_new = 0.0f, h_raw;
    if (COMP) {
      inu::add_compensated(hi, loi, -dti * sdh, h_new, lo_new);
      float s_raw, e_raw;
      inu::two_sum(hi, -dti * sdh, s_raw, e_raw);
      h_raw = s_raw + (e_raw + loi);
    } else {
      h_raw = FM ? fmaf(-dti, sdh, hi) : hi - dti * sdh;
      h_new = inu::gmax(h_raw, 0.0f);
    }
    float hu_new = FM ? fmaf(-dti, sdhu, hui) : hui - dti * sdhu;
    float hv_new = FM ? fmaf(-dti, sdhv, hvi) : hvi - dti * sdhv;
    if (REC) {
      // Adjoint tape of this stage: input, updated state before friction, clamp.
      const uint32_t slot = a.adj_cur->k - a.adj_cur->k0;
      float* t = a.tape + (size_t)(slot * kTapeWords + (ST == 0 ? 0u : 7u)) * a.N + i;
      const size_t q = a.N;
      t[0] = hi + loi;
      t[q] = hui;
      t[2 * q] = hvi;
      t[3 * q] = h_new + lo_new;
      t[4 * q] = hu_new;
      t[5 * q] = hv_new;
      t[6 * q] = (h_raw < 0.0f) ? 1.0f : 0.0f;
    }
    if (XTRA && a.track_clamp) {
      const float A = FM ? ldg(a.area, i) : aw;
      stg(a.clamp, i, ldg(a.clamp, i) + 0.5f * A * (h_new - h_raw));
    }
    if (h_new < p.dry) {
      hu_new = 0.0f;
      hv_new = 0.0f;
    } else if (FM) {
      inu::fast::friction(p, h_new, nm, dt, hu_new, hv_new);
    } else {
      const float drag = inu::film_drag(p, h_new, dt);
      hu_new *= drag;
      hv_new *= drag;
      const float u = hu_new / h_new;
      const float v = hv_new / h_new;
      const float spd = inu::sqrt_rn(u * u + v * v);
      const float h43 = inu::pow43(h_new);
      const float Sf = p.g * nm * nm * spd / h43;
      const float den = 1.0f + dt * Sf;
      hu_new = hu_new / den;
      hv_new = hv_new / den;
    }
    if (ST == 0) {
      stg(a.h1, i, h_new);
      if (COMP) stg(a.h1lo, i, lo_new);
      stg(a.hu1, i, hu_new);
      stg(a.hv1, i, hv_new);
    } else {
      float hf, lof = 0.0f;
      if (COMP) {
        float s, e;
        inu::two_sum(h0, h_new, s, e);
        e += lo0 + lo_new;
        inu::two_sum(0.5f * s, 0.5f * e, hf, lof);
      } else {
        hf = 0.5f * (h0 + h_new);
      }
      float huf = 0.5f * (hu0 + hu_new);
      float hvf = 0.5f * (hv0 + hv_new);
      if (hf < p.dry) {
        huf = 0.0f;
        hvf = 0.0f;
      }
      wet = hf >= p.dry;
      if (lam > 1e-30f) lam_b = lam;
      // Sources of this block (reference resident.py:695-708), after the
      // corrector; they change depth only.
      if (rate != 0.0f) {
        const float inc = dt * rate;
        if (COMP) {
          float hs, ls;
          inu::add_compensated(hf, lof, inc, hs, ls);
          hf = hs;
          lof = ls;
        } else {
          hf = inu::gmax(hf + inc, 0.0f);
        }
        wet = wet || hf >= p.dry;
      }
      stg(a.h, i, hf);
      if (COMP) stg(a.hlo, i, lof);
      stg(a.hu, i, huf);
      stg(a.hv, i, hvf);
      hi = hf;  // post-source depth, for the source dt bound below
      hui = huf;
      hvi = hvf;
    }
  }
  if (ST == 1) {
    const int n_wet = __syncthreads_count(wet ? 1 : 0);
    // Every cell at or above dry_tol after the sources. Sources that only
    // add water keep a wet cell wet (the compensated sum of a non-negative
    // increment never rounds below the pair's high word), so that is n_wet
    // == nc; a negative rate can dry a cell and needs its own count.
    const bool all_wet = a.src_neg ? __syncthreads_or((tid < nc && !(hi >= p.dry)) ? 1 : 0) == 0
                                   : (uint32_t)n_wet == nc;
    if (tid == 0) {
      a.wet[b] = n_wet != 0 ? 1u : 0u;
      a.part_dry[b] = all_wet ? 0u : 1u;
    }
    if (PERIM && a.n_src != 0u && all_wet) {
      // Every cell of the block is wet after the sources, so no neighbour in
      // it is dry and each edge term of the bound below is (u + c) len: the
      // sum is (u + c) times the perimeter. A dry neighbour in another block
      // is the source fix's, as below.
      if (tid < nc && rate != 0.0f) {
        float u, c;
        inu::fast::src_speeds(p, hi, hui, hvi, u, c);
        const float lam = (u + c) * perim;
        if (lam > 1e-30f) lam_b = inu::gmax(lam_b, lam);
      }
    } else if (a.n_src != 0u) {
      // Next-step dt bound of the block's wetted source cells
      // (resident.py:709-728) from the post-source depths, with every
      // neighbour in another block taken as wet; the source fix recomputes the
      // cells where such a neighbour is dry.
      if (tid < nc) l.cell[tid].x = hi;
      __syncthreads();
      if (tid < nc && rate != 0.0f && hi >= p.dry) {
        float u, c;
        src_uc<FM>(p, hi, hui, hvi, u, c);
        float lam = 0.0f;
        for (uint32_t k = k0; k < k1; ++k) {
          const uint32_t s = l.s_slot[k];
          const uint4 er = ldg(a.lrec, bi.e0 + (s >> 1));  // sides word, length (cached: read in phase A)
          const uint32_t q = (s & 1u) ? (er.x >> 16) : (er.x & 0xffffu);
          const bool dry_side = q < kBlockCells && l.cell[q].x < p.dry;  // walls and ghosts: wet
          const float kf = dry_side ? 2.0f : 1.0f;
          const float len = __uint_as_float(er.w);
          if (FM) lam = fmaf(fmaf(kf, c, u), len, lam);
          else lam += (u + kf * c) * len;
        }
        if (lam > 1e-30f) lam_b = inu::gmax(lam_b, lam);
      }
    }
    if (tid < nc) bound = cell_bound<FM>(p, aw, lam_b);
  }
  return bound;
}

// Loads of the walk-ahead state for the block at list position nx.t, all at
// once (the start of a walk, and blocks skipped without processing): its
// BlockInfo, this thread's ghost cell id, its scenario's dt and the list entry
// at nx.t2.
__device__ __forceinline__ void walk_ahead(const Args& a, Ahead& nx) {
  if (nx.t == kNoPos) return;
  nx.bi = ldc(a.binfo, nx.b);
  nx.gid = (threadIdx.x < nx.bi.n_ghost) ? a.gx[nx.bi.g0 + threadIdx.x] : 0u;
  nx.dtw = ldc(&a.scen[bi_scen(nx.bi)].dt_cur, 0u);
  if (nx.t2 != kNoPos) nx.b2 = ldc(a.list, nx.t2);
}

// Takes the next list position of the slice for the workgroup outside a
// block (a skipped block's turn): kNoPos once the slice is empty.
template <int ST>
__device__ __forceinline__ uint32_t walk_take(const Args& a, Walk& wk) {
  if (wk.last == kNoPos) return kNoPos;
  __syncthreads();  // every thread has read the word's last value
  if (threadIdx.x == 0u) s_walk_pos = atomicInc(walk_ctr<ST>(a), 0xffffffffu);
  __syncthreads();
  const uint32_t t = walk_pos<ST>(a, wk, __builtin_amdgcn_readfirstlane(s_walk_pos));
  if (t == kNoPos) wk.last = kNoPos;
  return t;
}

// Workgroups per CU the tile-0 stage kernels are compiled for: the predictor
// fits six in its registers; the corrector, which also holds the start state
// and sources of its cells across the edge phase, five without spilling.
template <int ST>
constexpr int stage_wgs() { return ST == 0 ? 6 : 5; }

// XTRA: the variant with the open-boundary pass and the clamp tally (each
// also switched at run time); the one without holds fewer values in scalar
// registers across the walk.
template <int ST, bool COMP, bool FM, int TILE, bool REC = false, bool ATOM = false, bool XTRA = true>
__global__ __launch_bounds__(256, TILE == 0 ? stage_wgs<ST>() : 1) void k_stage(const Args a) {
  // 16-byte aligned after the static LDS words: the records are read as float4.
  extern __shared__ __align__(16) float sm[];
  if (a.ctl->any_dt == 0u) return;  // no-op step: every scenario is at its stop
  const Lds l = carve_stage<ST>(sm, a.max_local, a.max_ghost);
  // XCD-aware split: workgroup w runs on XCD (w + r) % 8 for a rotation r of
  // the launch, so the workgroups of slice w % 8 share an XCD and walk one
  // contiguous (spatially compact) part of the block list, whose ghost cells
  // their L2 then keeps. A workgroup's first two blocks are fixed (list
  // positions j and j + per of the slice for the slice's j-th workgroup);
  // after those the blocks are taken in list order from a counter, one at a
  // time, so a workgroup that runs faster takes more of them and all of the
  // slice's workgroups end within about a block of each other.
  const uint32_t cnt = a.ctl->cnt_cur;
  const uint32_t xcd = blockIdx.x % kXcds;
  const uint32_t per = gridDim.x / kXcds;
  // The launchers size the grid in whole multiples of kXcds; a workgroup
  // past the last whole multiple would repeat a slice's first blocks and
  // upset the counter's final value.
  if (blockIdx.x >= per * kXcds) return;
  const uint32_t beg = (uint32_t)(((uint64_t)cnt * xcd) / kXcds);
  const uint32_t end = (uint32_t)(((uint64_t)cnt * (xcd + 1)) / kXcds);
  Walk wk;
  wk.base = beg + 2u * per;
  wk.len = end > wk.base ? end - wk.base : 0u;
  // Workgroups of the slice that take from its counter: those with a first
  // block. (None when the slice is empty; `last` is then never reached.)
  wk.last = wk.len + umin(per, end - beg) - 1u;
  Ahead nx;
  const uint32_t t0 = beg + blockIdx.x / kXcds;
  nx.t = t0 < end ? t0 : kNoPos;
  nx.t2 = t0 + per < end ? t0 + per : kNoPos;
  nx.b = 0u;
  nx.b2 = 0u;
  nx.gid = 0u;
  nx.dtw = 0u;
  if (nx.t != kNoPos) nx.b = ldc(a.list, nx.t);
  walk_ahead(a, nx);
  // This thread's smallest next-step bound, for scenario cur_s; flushed when
  // the walk moves to another scenario's blocks.
  uint32_t bound = kNoBound;
  uint32_t cur_s = (nx.t != kNoPos) ? bi_scen(nx.bi) : 0u;
  while (nx.t != kNoPos) {
    const uint32_t b = nx.b;
    const BlockInfo bi = nx.bi;
    const uint32_t gid = nx.gid;
    const uint32_t dtw = nx.dtw;
    nx.t = nx.t2;
    nx.b = __builtin_amdgcn_readfirstlane(nx.b2);
    if (ST == 1 && bi_scen(bi) != cur_s) {
      // Rare (a scenario boundary in this walk): flush per thread.
      if (bound != kNoBound && bound < __atomic_load_n(&a.scen[cur_s].dt_next, __ATOMIC_RELAXED))
        atomicMin(&a.scen[cur_s].dt_next, bound);
      bound = kNoBound;
      cur_s = bi_scen(bi);
    }
    if (dtw == 0u) {  // this scenario is at its stop
      nx.t2 = walk_take<ST>(a, wk);
      walk_ahead(a, nx);
      continue;
    }
    const uint32_t bb =
        stage_block<ST, COMP, FM, TILE, REC, ATOM, XTRA>(a, b, bi, gid, __uint_as_float(dtw), l, nx, wk);
    bound = (bb < bound) ? bb : bound;
    if (nx.t2 == kNoPos) wk.last = kNoPos;  // the block's take came back empty (or there was none)
  }
  if (ST == 1) block_atomic_min(&a.scen[cur_s].dt_next, bound);
}

// Sources of blocks the active set skipped this step (reference
// resident.py:691-708), listed by the dilation in `idle`: nothing else reads or
// writes those cells now. The block's post-source depths then give the dt
// bound of its wetted source cells (resident.py:709-728), with every
// neighbour in another block taken as wet, as in the corrector.
template <bool COMP, bool FM>
__global__ __launch_bounds__(256) void k_source_idle(const Args a) {
  __shared__ float s_h[kBlockCells];
  if (a.ctl->any_dt == 0u) return;
  const Phys& p = a.p;
  const uint32_t cnt = a.ctl->idle_cur;
  uint32_t bound = kNoBound;
  uint32_t cur_s = ~0u;  // scenario whose bound is being accumulated
  for (uint32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
    const uint32_t b = a.idle[t];
    const BlockInfo bi = a.binfo[b];
    if (bi_scen(bi) != cur_s) {
      if (cur_s != ~0u) block_atomic_min(&a.scen[cur_s].dt_next, bound);
      bound = kNoBound;
      cur_s = bi_scen(bi);
    }
    const uint32_t dtw = a.scen[cur_s].dt_cur;
    if (dtw == 0u) continue;  // this scenario is at its stop
    const float dt = __uint_as_float(dtw);
    const uint32_t c0 = b << 8;
    const uint32_t i = c0 + threadIdx.x;
    bool wet = false;
    float hn = 0.0f;
    if (threadIdx.x < bi_cells(bi)) {
      hn = a.h[i];
      const float rate = a.src[i];
      if (rate != 0.0f) {
        const float inc = dt * rate;
        if (COMP) {
          float ln;
          inu::add_compensated(hn, a.hlo[i], inc, hn, ln);
          a.hlo[i] = ln;
        } else {
          hn = inu::gmax(hn + inc, 0.0f);
        }
        a.h[i] = hn;
        wet = hn >= p.dry;
      }
    }
    __syncthreads();  // the previous block's s_h reads are done
    s_h[threadIdx.x] = hn;
    if (__syncthreads_or(wet ? 1 : 0) && threadIdx.x == 0) a.wet[b] = 1u;
    // Every cell of the block wet after the sources: the fast order takes
    // (u + c) times the perimeter, as the corrector does.
    const bool all_wet =
        __syncthreads_count((threadIdx.x >= bi_cells(bi) || hn >= p.dry) ? 1 : 0) == (int)blockDim.x;
    if (threadIdx.x == 0) a.part_dry[b] = all_wet ? 0u : 1u;
    if (wet) {
      const float hui = a.hu[i];
      const float hvi = a.hv[i];
      float u, c;
      src_uc<FM>(p, hn, hui, hvi, u, c);
      float lam = 0.0f, perim = 0.0f;
      const uint32_t s1 = a.cptr[i + 1];
      for (uint32_t s = a.cptr[i]; s < s1; ++s) {
        const uint32_t q = (uint32_t)a.snb[s] - c0;  // walls and other blocks fall outside
        const bool dry_side = q < kBlockCells && s_h[q] < p.dry;
        const float k = dry_side ? 2.0f : 1.0f;
        if (FM) lam = fmaf(fmaf(k, c, u), a.slen[s], lam);
        else lam += (u + k * c) * a.slen[s];
        perim += a.slen[s];
      }
      if (FM && all_wet) lam = (u + c) * perim;
      const uint32_t bb = cell_bound<FM>(p, FM ? a.inv_area[i] : a.area[i], lam);
      bound = (bb < bound) ? bb : bound;
    }
  }
  if (cur_s != ~0u) block_atomic_min(&a.scen[cur_s].dt_next, bound);
}

// Exact dt bound (reference resident.py:709-728) of the wetted source cells
// with a neighbour in another block that is dry after every source of the
// step: the corrector and k_source_idle counted that neighbour as wet. The
// bound with the dry side is never above the one taken before (every term
// and rounding is monotone), so the minimum over both equals the exact one.
// The pairs come in groups whose neighbours share a block; a group whose
// block is wet throughout (part_dry 0) has nothing to fix, which is most of
// them once the flood has spread. A workgroup takes kFixChunk consecutive
// pairs: it first reads the flags of the groups they touch, and only if one
// is set loads its pairs (kFixPer per thread, loads issued together),
// gathers the source cells whose neighbour is dry into LDS and bounds them
// a thread each, so a partly dry mesh keeps the parallelism of a thread per
// pair. Runs inside k_step_tail.
constexpr uint32_t kFixPer = 8;                    // pairs per thread
constexpr uint32_t kFixChunk = kFixPer * 256u;     // pairs per workgroup

// The exact bound of source cell i, whose neighbour in another block is dry.
template <bool FM>
__device__ __forceinline__ void fix_pair(const Args& a, uint32_t i) {
  const Phys& p = a.p;
  const float hn = a.h[i];
  if (!(hn >= p.dry) || a.src[i] == 0.0f) return;  // only wetted cells whose source is on
  const float hui = a.hu[i];
  const float hvi = a.hv[i];
  float u, c;
  src_uc<FM>(p, hn, hui, hvi, u, c);
  float lam = 0.0f;
  const uint32_t s1 = a.cptr[i + 1];
  for (uint32_t s = a.cptr[i]; s < s1; ++s) {
    const int nb = a.snb[s];
    const bool dry_side = nb >= 0 && a.h[nb] < p.dry;
    const float kf = dry_side ? 2.0f : 1.0f;
    if (FM) lam = fmaf(fmaf(kf, c, u), a.slen[s], lam);
    else lam += (u + kf * c) * a.slen[s];
  }
  const uint32_t bb = cell_bound<FM>(p, FM ? a.inv_area[i] : a.area[i], lam);
  if (bb != kNoBound) {
    // Rare: one atomic per such pair, into its scenario's slot. A scenario
    // at its stop ignores the slot (k_step_control).
    uint32_t* slot = &a.scen[bi_scen(a.binfo[i >> 8])].dt_next;
    if (bb < __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMin(slot, bb);
  }
}

template <bool FM>
__device__ __forceinline__ void source_fix_wg(const Args& a, uint32_t wg) {
  __shared__ uint32_t s_n;
  __shared__ uint32_t s_cell[kFixChunk];
  const uint2 gr = a.fix_chunk[wg];
  bool work = false;
  for (uint32_t g = gr.x + threadIdx.x; g < gr.y; g += blockDim.x) work = work || a.part_dry[a.fix_blk[g]] != 0u;
  if (threadIdx.x == 0) s_n = 0u;
  if (!__syncthreads_or(work ? 1 : 0)) return;
  const uint32_t k0 = wg * kFixChunk + threadIdx.x;
  uint2 pr[kFixPer];
#pragma unroll
  for (uint32_t u = 0; u < kFixPer; ++u) {
    const uint32_t k = k0 + 256u * u;
    pr[u] = (k < a.n_src_pairs) ? a.src_pair[k] : make_uint2(0u, 0u);
  }
  uint32_t pd[kFixPer];
  float hy[kFixPer];
#pragma unroll
  for (uint32_t u = 0; u < kFixPer; ++u) {
    pd[u] = a.part_dry[pr[u].y / kBlockCells];
    hy[u] = a.h[pr[u].y];
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (uint32_t u = 0; u < kFixPer; ++u) {
    const bool dry_nb = k0 + 256u * u < a.n_src_pairs && pd[u] != 0u && hy[u] < a.p.dry;
    const uint64_t m = __ballot(dry_nb ? 1 : 0);
    if (m != 0ull) {
      uint32_t base = 0u;
      if (lane == 0u) base = atomicAdd(&s_n, (uint32_t)__popcll(m));
      base = (uint32_t)__shfl((int)base, 0, 64);
      if (dry_nb) s_cell[base + (uint32_t)__popcll(m & below)] = pr[u].x;
    }
  }
  __syncthreads();
  const uint32_t n = s_n;
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) fix_pair<FM>(a, s_cell[q]);
}

// ---------------------------------------------------------------------------
// Active set and clock
// ---------------------------------------------------------------------------

// Start of a run: wet = a cell is wet or moving; or every block active with
// the active set off (reference resident.py:738-773). Also resets dt slot 0
// (cfl.py:130-166) and the list being built.
__global__ void k_block_mark(const Args a, int all_active) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < a.n_scen) a.scen[b].dt_cur = __float_as_uint(kDtSentinel);  // the prologue bound's slot
  if (b == 0) {
    a.ctl->cnt_next = 0u;
    a.ctl->idle_next = 0u;
  }
  if (b >= a.NB) return;
  if (all_active) {
    a.wet[b] = 0u;
    a.active[b] = 1u;
    return;
  }
  const uint32_t first = b << 8;
  const uint32_t last = first + bi_cells(a.binfo[b]);
  uint32_t live = 0u;
  for (uint32_t i = first; i < last; ++i) {
    if (a.h[i] >= a.p.dry || a.hu[i] != 0.0f || a.hv[i] != 0.0f) {
      live = 1u;
      break;
    }
  }
  a.wet[b] = live;
}

// Appends `v` of every lane with `take` to `out`, one atomic per wave.
__device__ __forceinline__ void wave_append(bool take, uint32_t v, uint32_t* cnt, uint32_t* out) {
  const uint64_t mask = __ballot(take ? 1 : 0);
  if (mask == 0ull) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, (int)leader, 64);
  if (take) out[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = v;
}

// active = (wet | forced) over the block's two-ring (reference
// resident.py:778-805). Active blocks are appended to `list`, inactive blocks
// holding a source to `idle`, in block order within each workgroup's
// kDilateBlocks blocks, with one atomic per workgroup and list (the
// counters are single words that every workgroup of the launch hits). The
// workgroup's {cells, owned edges, CSR slots} of the next step's active
// blocks go to act_part without atomics; k_step_control sums them. Blocks
// of a scenario that took a no-op step keep their flags and are left off
// both lists. The loads of a thread's blocks are issued together, and the
// two-ring is read eight neighbours at a time: the kernel is a chain of
// dependent loads, not a stream.
constexpr uint32_t kDilatePer = 2;                     // blocks per thread
constexpr uint32_t kDilateBlocks = kDilatePer * 256u;  // blocks per workgroup

__device__ __forceinline__ void dilate_wg(const Args& a, int force, int keep_flags, uint32_t wg) {
  __shared__ uint32_t s_off[2][kDilatePer * 4];  // per (round, wave): list, idle
  __shared__ uint32_t s_base[2];
  __shared__ uint32_t s_work[4][3];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t b0 = wg * kDilateBlocks + threadIdx.x;
  BlockInfo bi[kDilatePer];
  uint32_t v[kDilatePer], has[kDilatePer];
#pragma unroll
  for (uint32_t r = 0; r < kDilatePer; ++r) {
    const uint32_t b = b0 + r * 256u;
    if (b < a.NB) {
      bi[r] = a.binfo[b];
      v[r] = (keep_flags ? a.active[b] : (a.wet[b] | a.forced[b]));
      has[r] = a.has_src[b];
    } else {
      bi[r].meta = 0u;
      v[r] = has[r] = 0u;
    }
  }
  uint32_t w_cells = 0, w_edges = 0, w_slots = 0;
  uint64_t m_take[kDilatePer], m_idle[kDilatePer];
#pragma unroll
  for (uint32_t r = 0; r < kDilatePer; ++r) {
    const uint32_t b = b0 + r * 256u;
    bool take = false, take_idle = false;
    if (b < a.NB) {
      // A scenario at its stop took a no-op step: its flags stand. (Without
      // force the step moved some scenario; with one scenario, that one.)
      const bool frozen = !force && a.n_scen > 1u && a.scen[bi_scen(bi[r])].dt_cur == 0u;
      uint32_t act = v[r];
      if (frozen && !keep_flags) act = a.active[b];
      if (!keep_flags && !frozen) {
        const uint32_t k1 = a.bptr[b + 1];
        for (uint32_t k = a.bptr[b]; act == 0u && k < k1; k += 8u) {
          uint32_t n[8];
#pragma unroll
          for (uint32_t q = 0; q < 8u; ++q) n[q] = (k + q < k1) ? a.bidx[k + q] : b;
#pragma unroll
          for (uint32_t q = 0; q < 8u; ++q) act |= a.wet[n[q]] | a.forced[n[q]];
        }
        a.active[b] = act;
      }
      // A frozen scenario's blocks stay off the lists until a retarget
      // rebuilds them from the kept flags.
      take = act != 0u && !frozen;
      take_idle = act == 0u && !frozen && has[r] != 0u;
      if (take) {
        w_cells += bi_cells(bi[r]);
        w_edges += bi[r].n_own;
        w_slots += bi[r].n_slots;
      }
    }
    m_take[r] = __ballot(take ? 1 : 0);
    m_idle[r] = __ballot(take_idle ? 1 : 0);
    if (lane == 0u) {
      s_off[0][r * 4u + w] = (uint32_t)__popcll(m_take[r]);
      s_off[1][r * 4u + w] = (uint32_t)__popcll(m_idle[r]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    w_cells += (uint32_t)__shfl_xor((int)w_cells, off, 64);
    w_edges += (uint32_t)__shfl_xor((int)w_edges, off, 64);
    w_slots += (uint32_t)__shfl_xor((int)w_slots, off, 64);
  }
  if (lane == 0u) {
    s_work[w][0] = w_cells;
    s_work[w][1] = w_edges;
    s_work[w][2] = w_slots;
  }
  __syncthreads();
  if (threadIdx.x < 2u) {
    // Exclusive offsets in block order (round, then wave), one atomic.
    uint32_t run = 0;
    for (uint32_t j = 0; j < kDilatePer * 4u; ++j) {
      const uint32_t cnt = s_off[threadIdx.x][j];
      s_off[threadIdx.x][j] = run;
      run += cnt;
    }
    uint32_t* ctr = threadIdx.x == 0u ? &a.ctl->cnt_next : &a.ctl->idle_next;
    s_base[threadIdx.x] = run ? atomicAdd(ctr, run) : 0u;
  } else if (threadIdx.x >= 64u && threadIdx.x < 67u) {
    const uint32_t k = threadIdx.x - 64u;
    a.act_part[3u * wg + k] = (unsigned long long)s_work[0][k] + s_work[1][k] + s_work[2][k] + s_work[3][k];
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (uint32_t r = 0; r < kDilatePer; ++r) {
    const uint32_t b = b0 + r * 256u;
    if ((m_take[r] >> lane) & 1ull)
      a.list[s_base[0] + s_off[0][r * 4u + w] + (uint32_t)__popcll(m_take[r] & below)] = b;
    if ((m_idle[r] >> lane) & 1ull)
      a.idle[s_base[1] + s_off[1][r * 4u + w] + (uint32_t)__popcll(m_idle[r] & below)] = b;
  }
}

__global__ __launch_bounds__(256) void k_dilate(const Args a, int force, int keep_flags) {
  if (!force && a.ctl->any_dt == 0u) return;  // no-op step changed nothing
  dilate_wg(a, force, keep_flags, blockIdx.x);
}

// The end of a step before k_step_control, in one launch: workgroups
// [0, n_dil) run the dilation (dilate_wg), the rest the source fix
// (source_fix_wg). The two are independent of each other; both read what
// the corrector and k_source_idle wrote.
template <bool FM>
__global__ __launch_bounds__(256) void k_step_tail(const Args a, uint32_t n_dil) {
  if (a.ctl->any_dt == 0u) return;  // no-op step changed nothing
  if (blockIdx.x < n_dil) dilate_wg(a, 0, 0, blockIdx.x);
  else source_fix_wg<FM>(a, blockIdx.x - n_dil);
}

// Empties the lists being built (before a rebuild outside the step
// sequence).
__global__ void k_list_reset(const Args a) {
  a.ctl->cnt_next = 0u;
  a.ctl->idle_next = 0u;
}

// Prologue CFL bound of the current state, every cell (reference
// resident.py:528-610): per-edge wave bound into LDS, per-cell CSR sum,
// atomicMin into dt slot 0.
template <bool FM, bool LO>
__global__ __launch_bounds__(256) void k_prologue_cfl(const Args a) {
  extern __shared__ __align__(16) float sm[];
  const Lds l = carve(sm, a.max_local, a.max_ghost);
  const Phys& p = a.p;
  const uint32_t tid = threadIdx.x;
  uint32_t best = kNoBound;
  uint32_t cur_s = ~0u;  // scenario whose bound is being accumulated
  for (uint32_t b = blockIdx.x; b < a.NB; b += gridDim.x) {
    const uint32_t c0 = b << 8;
    const BlockInfo bi = a.binfo[b];
    const uint32_t nc = bi_cells(bi);
    if (bi_scen(bi) != cur_s) {
      if (cur_s != ~0u) block_atomic_min(&a.scen[cur_s].dt_cur, best);
      best = kNoBound;
      cur_s = bi_scen(bi);
    }
    __syncthreads();
    if (tid < nc)
      l.cell[2 * tid] = make_float4(a.h[c0 + tid], a.hu[c0 + tid], a.hv[c0 + tid], LO ? a.hlo[c0 + tid] : 0.0f);
    for (uint32_t g = tid; g < bi.n_ghost; g += kBlockCells)
      l.cell[2 * (kBlockCells + g)] = load_ghost<0, LO>(a, a.gx[bi.g0 + g]);
    __syncthreads();
    for (uint32_t j = tid; j < bi.n_loc; j += kBlockCells) {
      const EdgeRec r = load_edge(a, bi, j);
      const uint32_t ql = r.side >> 16;
      const uint32_t qr = r.side & 0xffffu;
      const bool wall = qr == kSideWall;
      const float4 sl = l.cell[2 * ql];
      float4 sr = wall ? sl : l.cell[2 * qr];
      float contrib = 0.0f;
      if (sl.x >= p.dry || sr.x >= p.dry) {
        if (FM) {
          const inu::fast::Cell L = inu::fast::cell(p, sl.x, sl.y, sl.z, sl.w);
          const inu::fast::Cell R = wall ? L : inu::fast::cell(p, sr.x, sr.y, sr.z, sr.w);
          contrib = inu::fast::wave_bound(p, L, R, wall, r.d, r.nx, r.ny) * r.len;
        } else {
          const float dzL = (r.d > 0.0f) ? r.d : 0.0f;
          const float dzR = (r.d < 0.0f) ? -r.d : 0.0f;
          if (wall) inu::wall_ghost(sl.y, sl.z, r.nx, r.ny, sr.y, sr.z);
          contrib = inu::wave_bound(p, sl.x, sl.y, sl.z, dzL, sr.x, sr.y, sr.z, dzR, r.nx, r.ny) * r.len;
        }
      }
      l.edge[2 * j].x = contrib;
    }
    __syncthreads();
    if (a.n_open != 0u) {
      // Open edges: the wave bound against the ghost of their condition
      // (reference resident.py:556-574).
      for (uint32_t t = a.open->bptr[b] + tid; t < a.open->bptr[b + 1]; t += kBlockCells) {
        const uint2 jk = a.open->jk[t];
        const EdgeRec r = load_edge(a, bi, jk.x);
        const float4 sl = l.cell[2 * (r.side >> 16)];
        float4 sr = sl;
        inu::ghost_state(p, a.open->kind[jk.y], a.open->value[jk.y], sl.x, sl.y, sl.z, r.nx, r.ny, sr.x, sr.y, sr.z);
        float contrib = 0.0f;
        if (sl.x >= p.dry || sr.x >= p.dry) {
          contrib = (FM ? inu::fast::wave_bound(p, inu::fast::cell(p, sl.x, sl.y, sl.z, sl.w),
                                                inu::fast::cell(p, sr.x, sr.y, sr.z, sl.w), false, 0.0f, r.nx, r.ny)
                        : inu::wave_bound(p, sl.x, sl.y, sl.z, 0.0f, sr.x, sr.y, sr.z, 0.0f, r.nx, r.ny)) *
                    r.len;
        }
        l.edge[2 * jk.x].x = contrib;
      }
      __syncthreads();
    }
    uint32_t bound = kNoBound;
    if (tid < nc) {
      const uint32_t i = c0 + tid;
      float lam = 0.0f;
      for (uint32_t k = a.cptr[i]; k < a.cptr[i + 1]; ++k) lam += l.edge[2 * (a.lidx[k] >> 1)].x;
      bound = cell_bound<FM>(p, FM ? a.inv_area[i] : a.area[i], lam);
    }
    best = (bound < best) ? bound : best;
  }
  if (cur_s != ~0u) block_atomic_min(&a.scen[cur_s].dt_cur, best);
}

// The integer clock of every scenario (reference resident.py:842-882), one
// thread per scenario. Mode 0 START plans from the prologue bound in dt_cur,
// mode 1 STEP advances past the step just taken and plans from dt_next, mode 2
// RETARGET re-plans from the last bound. START, RETARGET (after k_dilate
// rebuilt the lists from the kept flags) and every STEP that some scenario
// really took with dilation on publish the block lists k_dilate built.
__global__ void k_step_control(const Args a, int mode) {
  Ctl* t = a.ctl;
  bool real = false, moving = false;
  for (uint32_t s = threadIdx.x; s < a.n_scen; s += blockDim.x) {
    Scen& sc = a.scen[s];
    inu_ctrl& c = sc.ctrl;
    float raw;
    if (mode == 0) {
      raw = __uint_as_float(sc.dt_cur);
      c.last_raw = raw;
    } else if (mode == 1) {
      const uint32_t done = c.dt_ticks;
      c.now_ticks += (uint64_t)done;
      if (done > 0u) {
        c.steps += 1ull;
        raw = __uint_as_float(sc.dt_next);
        c.last_raw = raw;
        real = true;
      } else {
        raw = c.last_raw;
      }
    } else {
      raw = c.last_raw;
    }
    const uint64_t remaining = (c.stop_ticks > c.now_ticks) ? c.stop_ticks - c.now_ticks : 0ull;
    uint32_t ticks;
    if (remaining == 0ull || raw < 1e-10f) {
      ticks = 0u;
    } else if (raw >= 1e10f) {
      ticks = c.dry_dt_ticks;
    } else {
      const float cap = (float)c.dt_max_ticks / kTicksPerSecond;
      ticks = (uint32_t)(inu::gmin(raw * c.cfl_safety, cap) * kTicksPerSecond);
    }
    ticks = min(ticks, c.dt_max_ticks);
    if ((uint64_t)ticks > remaining) ticks = (uint32_t)remaining;
    c.dt_ticks = ticks;
    sc.dt_cur = __float_as_uint((float)ticks / kTicksPerSecond);
    sc.dt_next = __float_as_uint(kDtSentinel);
    moving = moving || ticks > 0u;
  }
  const int any_real = __syncthreads_or(real ? 1 : 0);
  const int any_moving = __syncthreads_or(moving ? 1 : 0);
  const bool swap = mode == 0 || mode == 2 || (mode == 1 && any_real && t->swap_on != 0u);
  // The next step's active-set work, summed over k_dilate's workgroups.
  __shared__ unsigned long long s_act[4][3];
  if (swap) {
    long part[3] = {0, 0, 0};
    for (uint32_t j = threadIdx.x; j < a.n_dwg; j += blockDim.x)
      for (int k = 0; k < 3; ++k) part[k] += (long)a.act_part[3u * j + k];
    for (int k = 0; k < 3; ++k)
      for (int off = 32; off > 0; off >>= 1) part[k] += __shfl_xor(part[k], off, 64);
    if ((threadIdx.x & 63u) == 0u)
      for (int k = 0; k < 3; ++k) s_act[threadIdx.x >> 6][k] = (unsigned long long)part[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 1 && any_real) {
      for (int k = 0; k < 3; ++k) t->act_total[k] += t->act_cur[k];
      t->act_steps += 1ull;
    }
    if (swap) {
      t->cnt_cur = t->cnt_next;
      t->cnt_next = 0u;
      t->idle_cur = t->idle_next;
      t->idle_next = 0u;
      for (int k = 0; k < 3; ++k) t->act_cur[k] = s_act[0][k] + s_act[1][k] + s_act[2][k] + s_act[3][k];
    }
    t->any_dt = any_moving ? 1u : 0u;
  }
}

// Volume through each open stretch after a step (reference resident.py:
// 613-658, run after each flux stage): each stage moves dt/2 times its mass
// word; out and in are kept apart, each as an exact pair of float32 words
// (hi, lo). One thread per stretch sums its edges in a fixed order, stage 0
// then stage 1; an edge of a block not visited this step moved nothing.
__global__ void k_open_tally(const Args a, const uint32_t* stretch_ptr, const uint32_t* open_cell, float* tally,
                             uint32_t n_stretch) {
  const uint32_t st = blockIdx.x * blockDim.x + threadIdx.x;
  if (st >= n_stretch) return;
  const uint32_t dtw = a.scen[0].dt_cur;
  if (dtw == 0u) return;
  const float half_dt = 0.5f * __uint_as_float(dtw);
  for (uint32_t stage = 0; stage < 2u; ++stage) {
    for (uint32_t k = stretch_ptr[st]; k < stretch_ptr[st + 1]; ++k) {
      if (a.active[open_cell[k] >> 8] == 0u) continue;
      const float v = half_dt * a.open->flux[(size_t)stage * a.n_open + k];
      const uint32_t o = 4u * st + ((v < 0.0f) ? 2u : 0u);
      float sum, err, hi, lo;
      inu::two_sum(tally[o], fabsf(v), sum, err);
      err += tally[o + 1];
      inu::two_sum(sum, err, hi, lo);
      tally[o] = hi;
      tally[o + 1] = lo;
    }
  }
}

// Staging of inu_set_source_rates_noalloc: up to kScratchRates internal cell
// ids, then as many rate words. Each GPU holds one copy for the life of the
// process; every call on a GPU shares it.
constexpr uint32_t kScratchRates = 1u << 16;
static __device__ __align__(256) uint32_t scratch[2 * kScratchRates];

// rates[k] into the dense source array at cells[k] (inu_set_source_rates).
__global__ void k_scatter_rates(float* src, const uint32_t* cells, const float* rates, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) src[cells[k]] = rates[k];
}

// ---------------------------------------------------------------------------
// Adjoint: reverse mode of recorded steps
// ---------------------------------------------------------------------------

// Arguments of one reverse stage. Adjoint arrays hold three arrays of N
// words each: d loss / d h, d hu, d hv. The step comes from the cursor: its dt
// from the log, its tape slot k - k0, its listed blocks from `list` (built by
// k_adj_lists).
struct AdjArgs {
  Args a;
  const float* tape;     // kTapeWords arrays of N per slot
  const uint32_t* log;   // dt bits per step
  AdjCur* cur;
  float* lam;            // d loss / d step output; d loss / d step input after stage 0
  float* l1;             // d loss / d stage-0 output (written by stage 1, read by stage 0)
  float* grad_n;         // += d loss / d Manning n, per cell
  float2* grad_dz;       // += d loss / d (dzL, dzR), per local edge stream position
};

// d loss / d corrector output of cell i from d loss / d step output (the Heun
// average, reference resident.py:489-501): half of each, the momenta's only
// where the averaged depth is wet (a dry average zeroes them). Sources are
// constant additions (non-negative rates, so the source clamp never cuts)
// and pass the depth gradient through. The same half is the step input's
// direct part.
__device__ __forceinline__ float3 heun_half(const float* tp, const float* lam, size_t N, uint32_t i, float dry) {
  const float h0 = tp[i], h2 = tp[10 * N + i];
  const bool wet = 0.5f * (h0 + h2) >= dry;
  return make_float3(0.5f * lam[i], wet ? 0.5f * lam[N + i] : 0.0f, wet ? 0.5f * lam[2 * N + i] : 0.0f);
}

// Reverse of one stage over the listed blocks (reference resident.py:401-516
// and :187-287 differentiated). Per block: (A) the friction reverse of its
// cells and of its listed ghost cells from the tape gives each one's update
// weight (-dt / area times d loss / d sums); a ghost whose block is not listed
// was not updated by this stage and weighs zero; (B) each local edge's word
// gradients from its two sides' weights, then the edge's VJP into LDS; the
// edge's bed-step gradient is kept by its owner or, when the owner is not
// listed, by this block; a nonzero gradient of a side outside the listed
// blocks (film there) is added to that cell's d loss / d step input at once,
// since such a cell keeps its state through the step; (C) each cell sums its
// edges' contributions in CSR order. Deterministic except for those
// boundary additions, which are counted.
template <int ST>
__global__ __launch_bounds__(256) void k_adj_stage(const AdjArgs x) {
  extern __shared__ __align__(16) float sm[];
  const Args& a = x.a;
  const Lds l = carve(sm, a.max_local, a.max_ghost);
  const Phys& p = a.p;
  const uint32_t tid = threadIdx.x;
  const size_t N = a.N;
  const uint32_t k = x.cur->k;
  const float* slot = x.tape + (size_t)(k - x.cur->k0) * kTapeWords * N;
  const float* tp = slot + (ST == 0 ? 0u : 7u) * N;
  const float dt = __uint_as_float(x.log[k]);
  const float* lam_out = (ST == 0) ? x.l1 : nullptr;
  uint32_t n_cross = 0;
  // The stage kernels' walk (k_stage): a slice of the list per XCD, each
  // workgroup's first block fixed, then one block at a time from the
  // slice's self-clearing counter (taken while the block is reversed).
  __shared__ uint32_t s_adj_pos;
  const uint32_t cnt = a.ctl->cnt_cur;
  const uint32_t per = gridDim.x / kXcds;
  if (blockIdx.x >= per * kXcds) return;  // the host sizes the grid in whole multiples
  const uint32_t xcd = blockIdx.x % kXcds;
  const uint32_t beg = (uint32_t)(((uint64_t)cnt * xcd) / kXcds);
  const uint32_t end = (uint32_t)(((uint64_t)cnt * (xcd + 1)) / kXcds);
  uint32_t* const ctr = a.walk + ((2u + (uint32_t)ST) * kXcds + xcd) * kWalkStride;
  Walk wk;
  wk.base = beg + per;
  wk.len = end > wk.base ? end - wk.base : 0u;
  wk.last = wk.len + umin(per, end - beg) - 1u;
  uint32_t t = beg + blockIdx.x / kXcds;
  if (t >= end) t = kNoPos;
  while (t != kNoPos) {
    uint32_t taken = kNoPos;
    if (tid == 0u && wk.last != kNoPos) taken = atomicInc(ctr, 0xffffffffu);
    const uint32_t b = a.list[t];
    const BlockInfo bi = a.binfo[b];
    const uint32_t c0 = b << 8, nc = bi_cells(bi);
    float rh = 0.0f, rhu = 0.0f, rhv = 0.0f;
    __syncthreads();  // the previous block's LDS reads are done
    // ---- (A) cells and ghosts -----------------------------------------------
    for (uint32_t q = tid; q < kBlockCells + bi.n_ghost; q += kBlockCells) {
      const bool own = q < kBlockCells;
      if (own && q >= nc) continue;
      const uint32_t i = own ? c0 + q : a.gx[bi.g0 + (q - kBlockCells)];
      const bool listed = own || a.active[i >> 8] != 0u;
      const float hin = tp[i], huin = tp[N + i], hvin = tp[2 * N + i];
      float gh = 0.0f, ghu = 0.0f, ghv = 0.0f;
      if (listed) {
        const float hp = tp[3 * N + i], hup = tp[4 * N + i], hvp = tp[5 * N + i];
        const bool clamped = tp[6 * N + i] != 0.0f;
        const float3 lo = (ST == 1) ? heun_half(slot, x.lam, N, i, p.dry)
                                    : make_float3(lam_out[i], lam_out[N + i], lam_out[2 * N + i]);
        float gn;
        inu::adj::friction_vjp(p, hp, hup, hvp, clamped, a.nman[i], dt, lo.x, lo.y, lo.z, gh, ghu, ghv, gn);
        if (own) {
          rh = gh;
          rhu = ghu;
          rhv = ghv;
          x.grad_n[i] += gn;
        }
      }
      const float dti = dt / a.area[i];
      l.cell[2 * q] = make_float4(hin, huin, hvin, listed ? 1.0f : 0.0f);
      l.cell[2 * q + 1] = make_float4(-dti * gh, -dti * ghu, -dti * ghv, 0.0f);
    }
    __syncthreads();
    // ---- (B) edges -------------------------------------------------------------
    for (uint32_t j = tid; j < bi.n_loc; j += kBlockCells) {
      const EdgeRec r = load_edge(a, bi, j);
      const uint32_t ql = r.side >> 16, qr = r.side & 0xffffu;
      const bool wall = qr == kSideWall;
      const float4 sl = l.cell[2 * ql], bl = l.cell[2 * ql + 1];
      const float4 sr = wall ? sl : l.cell[2 * qr];
      const float4 br = wall ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : l.cell[2 * qr + 1];
      // w0 enters the left sum and leaves the right one; w1, w2 are the
      // left momentum words, w3, w4 the right ones (subtracted).
      const float lw[5] = {bl.x - br.x, bl.y, bl.z, -br.y, -br.z};
      const float dzL = (r.d > 0.0f) ? r.d : 0.0f;
      const float dzR = (r.d < 0.0f) ? -r.d : 0.0f;
      float gL[3], gR[3], gdz[2];
      inu::adj::edge_vjp(p, wall, sl.x, sl.y, sl.z, sr.x, sr.y, sr.z, dzL, dzR, r.nx, r.ny, r.len, lw, gL, gR, gdz);
      l.edge[2 * j] = make_float4(gL[0], gL[1], gL[2], 0.0f);
      l.edge[2 * j + 1] = make_float4(gR[0], gR[1], gR[2], 0.0f);
      // Edges with only the right cell here are owned by the left cell's block.
      const bool keep_dz = j < bi.n_own || sl.w == 0.0f;
      if (keep_dz && !wall) {
        float2 acc = x.grad_dz[bi.e0 + j];
        acc.x += gdz[0];
        acc.y += gdz[1];
        x.grad_dz[bi.e0 + j] = acc;
      }
      // A side outside the listed blocks is a ghost cell (own cells are listed).
      if (sl.w == 0.0f && (gL[0] != 0.0f || gL[1] != 0.0f || gL[2] != 0.0f)) {
        const uint32_t i = a.gx[bi.g0 + (ql - kBlockCells)];
        atomicAdd(&x.lam[i], gL[0]);
        atomicAdd(&x.lam[N + i], gL[1]);
        atomicAdd(&x.lam[2 * N + i], gL[2]);
        ++n_cross;
      }
      if (!wall && sr.w == 0.0f && (gR[0] != 0.0f || gR[1] != 0.0f || gR[2] != 0.0f)) {
        const uint32_t i = a.gx[bi.g0 + (qr - kBlockCells)];
        atomicAdd(&x.lam[i], gR[0]);
        atomicAdd(&x.lam[N + i], gR[1]);
        atomicAdd(&x.lam[2 * N + i], gR[2]);
        ++n_cross;
      }
    }
    __syncthreads();
    // ---- (C) cells ---------------------------------------------------------------
    if (tid < nc) {
      const uint32_t i = c0 + tid;
      float sh = rh, shu = rhu, shv = rhv;
      const uint32_t k1 = a.cptr[i + 1];
      for (uint32_t kk = a.cptr[i]; kk < k1; ++kk) {
        const uint32_t sidx = a.lidx[kk];
        const float4 g = l.edge[2 * (sidx >> 1) + (sidx & 1u)];
        sh += g.x;
        shu += g.y;
        shv += g.z;
      }
      if (ST == 0) {
        // The step input's direct part, from the Heun average.
        const float3 d = heun_half(slot, x.lam, N, i, p.dry);
        sh += d.x;
        shu += d.y;
        shv += d.z;
        x.lam[i] = sh;
        x.lam[N + i] = shu;
        x.lam[2 * N + i] = shv;
      } else {
        x.l1[i] = sh;
        x.l1[N + i] = shu;
        x.l1[2 * N + i] = shv;
      }
    }
    // The next block: the position taken above (its slice's final take
    // clears the counter).
    if (tid == 0u) s_adj_pos = taken;
    __syncthreads();
    const uint32_t r = __builtin_amdgcn_readfirstlane(s_adj_pos);
    if (tid == 0u && r != kNoPos && r == wk.last) atomicExch(ctr, 0u);
    t = r < wk.len ? wk.base + r : kNoPos;
    if (t == kNoPos) wk.last = kNoPos;
  }
  if (n_cross) atomicAdd(&x.cur->stats[2], (unsigned long long)n_cross);
}

// What a recording keeps per step (k_adj_record).
struct AdjRec {
  uint32_t* log;       // dt bits per step
  uint8_t* codes;      // per step and block: 1 listed, 2 idle source block, 0 neither
  float* series;       // gauge depths, one row per sample
  const uint32_t* gauges;
  float* ck;           // state checkpoints (h, hu, hv, h_lo), one every seg steps
  uint32_t max_steps, n_g, every, max_samples, seg, n_ck;
};

// Recording, at the start and after every step (k = steps taken): the next
// step's dt and its lists (the rows start zeroed), gauge depths every
// `every` steps, the state every `seg` steps. A no-op step rewrites row k
// with its own (empty) lists, which no replay reads.
__global__ __launch_bounds__(256) void k_adj_record(const Args a, const AdjRec r) {
  const uint64_t k = a.scen[0].ctrl.steps;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (k < r.max_steps) {
    if (tid == 0) r.log[k] = a.scen[0].dt_cur;
    uint8_t* row = r.codes + k * a.NB;
    const uint32_t na = a.ctl->cnt_cur, ni = a.ctl->idle_cur;
    for (uint32_t t = tid; t < na; t += stride) row[a.list[t]] = 1u;
    for (uint32_t t = tid; t < ni; t += stride) row[a.idle[t]] = 2u;
  }
  if (k > 0 && k % r.every == 0 && k / r.every <= r.max_samples)
    for (uint32_t j = tid; j < r.n_g; j += stride) {
      const uint32_t g = r.gauges[j];
      r.series[(k / r.every - 1) * r.n_g + j] = a.h[g] + a.hlo[g];
    }
  if (k % r.seg == 0 && k / r.seg < r.n_ck) {
    float* dst = r.ck + (k / r.seg) * 4ull * a.N;
    for (size_t i = tid; i < a.N; i += stride) {
      dst[i] = a.h[i];
      dst[a.N + i] = a.hu[i];
      dst[2 * (size_t)a.N + i] = a.hv[i];
      dst[3 * (size_t)a.N + i] = a.hlo[i];
    }
  }
}

// Replay and reverse of step k = cursor: its dt, its active flags and lists
// from the recorded codes (the counts start at zero), and, in the reverse
// pass (seed != null), d loss / d the gauge depths sampled after the step
// (gauges are distinct).
struct AdjStep {
  const uint8_t* codes;
  const uint32_t* log;
  const AdjCur* cur;
  const float* seed;  // per sample row, or null
  const uint32_t* gauges;
  float* lam;
  uint32_t n_g, every, max_samples;
};

__global__ __launch_bounds__(256) void k_adj_lists(const Args a, const AdjStep s) {
  const uint32_t k = s.cur->k;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t code = 0u;
  if (b < a.NB) {
    code = s.codes[(size_t)k * a.NB + b];
    a.active[b] = code == 1u ? 1u : 0u;
  }
  wave_append(code == 1u, b, &a.ctl->cnt_cur, a.list);
  wave_append(code == 2u, b, &a.ctl->idle_cur, a.idle);
  if (b == 0) {
    a.scen[0].dt_cur = s.log[k];
    a.ctl->any_dt = 1u;
  }
  if (s.seed && (k + 1) % s.every == 0 && (k + 1) / s.every <= s.max_samples) {
    const float* row = s.seed + (size_t)((k + 1) / s.every - 1) * s.n_g;
    for (uint32_t j = b; j < s.n_g; j += gridDim.x * blockDim.x) s.lam[s.gauges[j]] += row[j];
  }
}

// End of a replayed (dir = +1) or reversed (dir = -1) step: the replay's
// gauge sample (series != null), the reverse pass's list counters, the
// counts cleared for the next k_adj_lists, the cursor moved.
__global__ __launch_bounds__(256) void k_adj_end(const Args a, AdjCur* cur, int dir, float* series,
                                                 const uint32_t* gauges, uint32_t n_g, uint32_t every,
                                                 uint32_t max_samples) {
  const uint32_t k = cur->k;
  if (series && (k + 1) % every == 0 && (k + 1) / every <= max_samples)
    for (uint32_t j = threadIdx.x; j < n_g; j += blockDim.x) {
      const uint32_t g = gauges[j];
      series[(size_t)((k + 1) / every - 1) * n_g + j] = a.h[g] + a.hlo[g];
    }
  __syncthreads();  // every thread has read the cursor
  if (threadIdx.x == 0) {
    if (dir < 0) {
      cur->stats[0] += a.ctl->cnt_cur;
      cur->stats[1] += a.ctl->idle_cur;
    }
    a.ctl->cnt_cur = 0u;
    a.ctl->idle_cur = 0u;
    cur->k = k + (uint32_t)dir;
  }
}

// Cursor at step k of the segment starting at k0, list counts cleared.
__global__ void k_adj_cursor(const Args a, AdjCur* cur, uint32_t k, uint32_t k0) {
  cur->k = k;
  cur->k0 = k0;
  a.ctl->cnt_cur = 0u;
  a.ctl->idle_cur = 0u;
}

// Edge-word VJP probe: per row in17 = {hL, huL, hvL, hR, huR, hvR, dzL, dzR,
// nx, ny, len, wall, lw0..lw4}, out8 = {gL, gR, gdzL, gdzR}.
__global__ void k_probe_vjp(const float* in, float* out, uint32_t n, Phys p) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float* r = in + 17ull * k;
  float gL[3], gR[3], gdz[2];
  inu::adj::edge_vjp(p, r[11] != 0.0f, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r + 12, gL,
                     gR, gdz);
  float* o = out + 8ull * k;
  o[0] = gL[0];
  o[1] = gL[1];
  o[2] = gL[2];
  o[3] = gR[0];
  o[4] = gR[1];
  o[5] = gR[2];
  o[6] = gdz[0];
  o[7] = gdz[1];
}

// Plain copy of n 16-byte words: the kernel of known traffic that HBM byte
// counters are scaled against.
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ in, float4* __restrict__ out, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    out[k] = in[k];
}

// Bandwidth probes of the attainable HBM rate (measurement only). Each thread
// moves `U` 16-byte words per trip, all loads issued before any store, with
// nontemporal hints (streamed data is not kept in the caches); the grid is a
// whole number of workgroups per CU.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int U>
__global__ __launch_bounds__(256) void k_bw_copy(const f32x4* __restrict__ in, f32x4* __restrict__ out, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256u * U;
  for (size_t base = (size_t)blockIdx.x * 256u * U + threadIdx.x; base < n; base += stride) {
    f32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + (size_t)u * 256u < n) v[u] = __builtin_nontemporal_load(in + base + (size_t)u * 256u);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + (size_t)u * 256u < n) __builtin_nontemporal_store(v[u], out + base + (size_t)u * 256u);
  }
}

// Read only: every word summed, one word per thread written at the end.
template <int U>
__global__ __launch_bounds__(256) void k_bw_read(const f32x4* __restrict__ in, float* __restrict__ out, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256u * U;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (size_t base = (size_t)blockIdx.x * 256u * U + threadIdx.x; base < n; base += stride) {
    f32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = (base + (size_t)u * 256u < n) ? __builtin_nontemporal_load(in + base + (size_t)u * 256u)
                                            : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
  }
  out[(size_t)blockIdx.x * 256u + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

// Write only.
template <int U>
__global__ __launch_bounds__(256) void k_bw_write(f32x4* __restrict__ out, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256u * U;
  const f32x4 v = {1.0f, 2.0f, 3.0f, 4.0f};
  for (size_t base = (size_t)blockIdx.x * 256u * U + threadIdx.x; base < n; base += stride) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + (size_t)u * 256u < n) __builtin_nontemporal_store(v, out + base + (size_t)u * 256u);
  }
}

// Arithmetic probes: pairs (a, b) -> two-sum error, a / b, sqrt|a|, a^(4/3),
// film drag of depth a at dt b, a^(-1/3), fast film drag (reference
// solver.py:449-467 checks that the two-sum error term survives compilation).
__global__ void k_probe(const float* in, float* out, uint32_t n, Phys p) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float x = in[2 * k], y = in[2 * k + 1];
  float s, e;
  inu::two_sum(x, y, s, e);
  float* o = out + (size_t)kProbeOut * k;
  o[0] = e;
  o[1] = x / y;
  o[2] = inu::sqrt_rn(fabsf(x));
  o[3] = (x > 0.0f) ? inu::pow43(x) : 0.0f;
  o[4] = (x > 0.0f) ? inu::film_drag(p, x, y) : 0.0f;
  o[5] = (x > 0.0f) ? inu::fast::rcbrt(x) : 0.0f;
  o[6] = (x >= p.dry && y >= 0.0f) ? inu::fast::film_drag(p, x, y) : 0.0f;  // its domain
}

// Reciprocal probe: every float bit pattern u in [lo, hi), and its negative,
// compared bitwise between the fast order's rcp and the IEEE division 1 / x
// (two NaNs count as equal). counts[0] gets the number that differ, first[0]
// the smallest such pattern.
__global__ void k_probe_rcp(uint32_t lo, uint32_t hi, unsigned long long* counts, uint32_t* first) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (uint64_t)(hi - lo);
       k += (uint64_t)gridDim.x * blockDim.x) {
    for (uint32_t sign = 0; sign < 2; ++sign) {
      const uint32_t u = (lo + (uint32_t)k) | (sign << 31);
      const float x = __uint_as_float(u);
      const float f = inu::fast::rcp(x);
      const float q = 1.0f / x;
      if (__float_as_uint(f) != __float_as_uint(q) && !(f != f && q != q)) {
        atomicAdd(counts, 1ull);
        atomicMin(first, u);
      }
    }
  }
}

// Open boundary probe: cases (h, hu, hv, nx, ny, fh) with a condition and
// value each -> the ghost state (hG, huG, hvG) and the boundary mass flux of
// an HLLC mass flux fh.
__global__ void k_probe_open(const float* in, const uint32_t* kind, const float* value, float* out, uint32_t n,
                             Phys p) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float* c = in + 6 * (size_t)k;
  float* o = out + 4 * (size_t)k;
  inu::ghost_state(p, kind[k], value[k], c[0], c[1], c[2], c[3], c[4], o[0], o[1], o[2]);
  o[3] = inu::boundary_mass(kind[k], value[k], c[5]);
}

}  // namespace

// ===========================================================================
// Host runtime
// ===========================================================================

struct inu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cus = 256;
  // mesh
  uint32_t N = 0, E = 0, S = 0, NB = 0, max_local = 0, n_halo = 0;
  bool has_mesh = false;
  // params
  Phys p{};
  uint32_t flags = INU_FLAG_COMPENSATED | INU_FLAG_ACTIVE_SET;
  // scenarios: caller cells [cell0[s], cell0[s + 1]), blocks [block0[s], block0[s + 1])
  uint32_t n_scen = 0;
  std::vector<uint32_t> cell0, block0;
  Scen* scen = nullptr;
  // sources
  uint32_t n_src = 0, n_src_blocks = 0, n_src_pairs = 0;
  std::vector<uint32_t> src_cells;  // internal ids of the source cells, ascending
  // device memory
  std::vector<void*> allocs;
  Args a{};
  Ctl* ctl = nullptr;
  // launch shape
  uint32_t grid_stage = 0;        // default corrector grid (reported by inu_mesh_info)
  uint32_t grid_prologue = 0;     // k_prologue_cfl's grid
  uint32_t grid[2][2][2][kTiles] = {};  // [stage][compensated][fast order][tile]
  int tile = 0;                         // stage kernel tile level (inu_autotune)
  size_t lds_stage0 = 0, lds_stage1 = 0;  // the predictor's and the corrector's
  size_t lds_full = 0;                     // the prologue's and the adjoint's (carve())
  // (cell, neighbour) for every neighbour in another block, by cell
  std::vector<std::pair<uint32_t, uint32_t>> cross;
  // Edges (union numbering), for the bed gradient and inu_set_bed: cells
  // (internal ids, -1 wall), each stream position's edge, each edge's owner
  // position, and the signed bed steps in stream order.
  std::vector<int32_t> e_cl, e_cr;
  std::vector<uint32_t> stream_edge, own_pos;
  std::vector<float> h_ldz;
  bool src_nonneg = true;  // every source rate so far is >= 0
  // adjoint recording
  struct {
    bool on = false;
    uint32_t max_steps = 0, seg = 1, every = 1, n_gauges = 0, max_samples = 0, n_ck = 0;
    // 0: a checkpoint every segment, written by the recording; else the
    // number of checkpoint slots of the binomial schedule (slot 0 holds the
    // recording's start, the others are filled by the backward pass).
    uint32_t slots = 0;
    // steps the last backward replayed without tape, and with tape
    unsigned long long replayed[2] = {0ull, 0ull};
    uint32_t* log = nullptr;
    uint8_t* codes = nullptr;  // per step and block: 1 listed, 2 idle source block
    uint32_t* gauges = nullptr;
    float* series = nullptr;
    float* ck = nullptr;
    AdjCur* cur = nullptr;
    std::vector<std::pair<uint32_t, float*>> events;  // (step, dense source rates from that step)
    // Work of the replay and reverse passes, kept with the recording (its
    // graphs point at them): tape of one segment, d loss / d state, d loss /
    // d stage-0 output, parameter gradients, gauge seeds, replayed series.
    float *tape = nullptr, *lam = nullptr, *l1 = nullptr, *gn = nullptr, *dseed = nullptr, *rseries = nullptr;
    float2* gdz = nullptr;
    uint32_t grid_adj = 1;
    unsigned long long stats[3] = {0ull, 0ull, 0ull};
  } adj;
  // Reverse pass: face depth below which a side's velocity and celerity are
  // held; negative: kAdjThinDry x dry_tol.
  float adj_thin = -1.0f;
  // Checkpoint slots of the next recording's binomial schedule (0: uniform).
  uint32_t adj_slots = 0;
  // Open boundaries: open edge k's inside cell (internal id), the stretches
  // (open edges [ptr[s], ptr[s + 1]) form stretch s) and their volume tally.
  uint32_t n_open = 0, n_stretch = 0;
  std::vector<uint32_t> open_cell_h;
  uint32_t *open_cell = nullptr, *stretch_ptr = nullptr;
  float* tally = nullptr;
  float4* open_geom = nullptr;
  OpenDev open_h{};  // host copy of the device record's pointers
  // graphs of recorded step blocks, by step count
  std::map<uint32_t, hipGraphExec_t> graphs;
  std::vector<hipEvent_t> events;
};

namespace {

int fail(inu_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return fail((ctx), INU_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

template <typename T>
int dalloc(inu_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  void* q = nullptr;
  HIPCHK(c, hipMalloc(&q, n * sizeof(T)));
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return 0;
}

// New depths from the host: every block may hold a dry cell until a step
// has looked (the source fix reads part_dry).
int depths_replaced(inu_ctx* c) {
  HIPCHK(c, hipMemsetAsync(c->a.part_dry, 1, (size_t)c->NB * 4, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <typename T>
int upload(inu_ctx* c, T* dst, const T* src, size_t n) {
  if (n == 0) return 0;
  HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Caller cell arrays (scenarios back to back) to and from the internal layout
// (each scenario from a block boundary).
int upload_cells(inu_ctx* c, float* dev, const float* host) {
  for (uint32_t s = 0; s < c->n_scen; ++s) {
    const size_t n = c->cell0[s + 1] - c->cell0[s];
    if (n) HIPCHK(c, hipMemcpyAsync(dev + (size_t)c->block0[s] * kBlockCells, host + c->cell0[s], n * 4,
                                    hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int download_cells(inu_ctx* c, const float* dev, float* host) {
  for (uint32_t s = 0; s < c->n_scen; ++s) {
    const size_t n = c->cell0[s + 1] - c->cell0[s];
    if (n) HIPCHK(c, hipMemcpyAsync(host + c->cell0[s], dev + (size_t)c->block0[s] * kBlockCells, n * 4,
                                    hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Internal cell of caller cell u (u < N).
uint32_t internal_cell(const inu_ctx* c, uint32_t u) {
  const size_t s = (size_t)(std::upper_bound(c->cell0.begin(), c->cell0.end(), u) - c->cell0.begin()) - 1;
  return c->block0[s] * kBlockCells + (u - c->cell0[s]);
}

void free_all(inu_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  for (void* q : c->allocs) (void)hipFree(q);
  c->allocs.clear();
  c->a = Args{};
  c->ctl = nullptr;
  c->scen = nullptr;
  // The recording's buffers went with the rest: forget their addresses, which
  // the next allocations may reuse (releasing the recording later must not
  // free those).
  auto& d = c->adj;
  d.on = false;
  d.events.clear();
  d.log = nullptr;
  d.codes = nullptr;
  d.gauges = nullptr;
  d.series = nullptr;
  d.ck = nullptr;
  d.cur = nullptr;
  d.tape = d.lam = d.l1 = d.gn = d.dseed = d.rseries = nullptr;
  d.gdz = nullptr;
  c->n_scen = 0;
  c->has_mesh = false;
}

Phys make_phys(float g, float dry, float cfl) {
  Phys p;
  p.g = g;
  p.dry = dry;
  p.cfl = cfl;
  volatile float fd = 4.0f;  // FILM_DEPTHS
  p.e_film = fd * dry;
  p.e_film2 = p.e_film * p.e_film;
  p.half_g = 0.5f * g;
  p.sqrt_g_e = std::sqrt(g / p.e_film);
  p.inv_dry = 1.0f / dry;
  p.sqrt_g = std::sqrt(g);
  p.adj_thin = 0.0f;
  return p;
}

int check_ready(inu_ctx* c) {
  if (!c) return INU_EINVAL;
  if (!c->has_mesh) return fail(c, INU_ESTATE, "no mesh: call inu_set_mesh first");
  return 0;
}

// A copy on the solver stream, waited for: ordered after every kernel and
// copy already queued there.
int copy_sync(inu_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return 0;
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, kind, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// The scenarios' records (control block and dt words), copied whole.
int scen_read(inu_ctx* c, std::vector<Scen>& out) {
  out.resize(c->n_scen);
  return copy_sync(c, out.data(), c->scen, sizeof(Scen) * c->n_scen, hipMemcpyDeviceToHost);
}

// Kernel launches of one step, in stream order. With `ev` set, events are
// recorded before the predictor, between the two stages and after the
// corrector (no host wait).
// Launches a stage kernel variant: the run-time flags pick the compensated,
// fast-order and XTRA instantiations.
template <int ST, bool CP, bool F, int TILE, bool REC, bool ATOM>
void launch_k_stage(inu_ctx* c, const Args& a, dim3 g, size_t lds) {
  if (a.n_open != 0u || a.track_clamp != 0u)
    hipLaunchKernelGGL((k_stage<ST, CP, F, TILE, REC, ATOM, true>), g, dim3(256), lds, c->stream, a);
  else
    hipLaunchKernelGGL((k_stage<ST, CP, F, TILE, REC, ATOM, false>), g, dim3(256), lds, c->stream, a);
}
template <int ST, int TILE, bool REC, bool ATOM>
void launch_k_stage(inu_ctx* c, const Args& a, bool comp, bool fm, dim3 g, size_t lds) {
  if (comp && fm) launch_k_stage<ST, true, true, TILE, REC, ATOM>(c, a, g, lds);
  else if (comp) launch_k_stage<ST, true, false, TILE, REC, ATOM>(c, a, g, lds);
  else if (fm) launch_k_stage<ST, false, true, TILE, REC, ATOM>(c, a, g, lds);
  else launch_k_stage<ST, false, false, TILE, REC, ATOM>(c, a, g, lds);
}

template <int ST, int TILE>
void launch_stage_t(inu_ctx* c, const Args& a, bool comp, bool fm) {
  const size_t lds = ST == 0 ? c->lds_stage0 : c->lds_stage1;
  const dim3 g(c->grid[ST][comp ? 1 : 0][fm ? 1 : 0][TILE]);
  launch_k_stage<ST, TILE, false, false>(c, a, comp, fm, g, lds);
}

// Unordered accumulation (INU_FLAG_DETERMINISTIC off), tile level 0.
template <int ST>
void launch_stage_atom(inu_ctx* c, const Args& a, bool comp, bool fm) {
  // The four per-cell sums (4 x 256 floats) take the edge records' place.
  const size_t lds = std::max(ST == 0 ? c->lds_stage0 : c->lds_stage1,
                              (6 * (size_t)stage_cells(a.max_ghost) + 4 * (size_t)kBlockCells) * sizeof(float));
  const dim3 g(c->grid[ST][comp ? 1 : 0][fm ? 1 : 0][0]);
  launch_k_stage<ST, 0, false, true>(c, a, comp, fm, g, lds);
}

template <int ST>
void launch_stage(inu_ctx* c, const Args& a, bool comp, bool fm, int tile) {
  if (!(c->flags & INU_FLAG_DETERMINISTIC)) launch_stage_atom<ST>(c, a, comp, fm);
  else if (tile == 0) launch_stage_t<ST, 0>(c, a, comp, fm);
  else if (tile == 1) launch_stage_t<ST, 1>(c, a, comp, fm);
  else launch_stage_t<ST, 2>(c, a, comp, fm);
}

template <int ST>
void launch_stage(inu_ctx* c, bool comp, bool fm) {
  launch_stage<ST>(c, c->a, comp, fm, c->tile);
}

// Sources of the blocks listed in `idle` (k_source_idle).
void launch_source_idle(inu_ctx* c, const Args& a, bool comp, bool fm) {
  const uint32_t gb = std::max(1u, std::min<uint32_t>(c->n_src_blocks, 8u * (uint32_t)c->cus));
  if (comp && fm) hipLaunchKernelGGL((k_source_idle<true, true>), dim3(gb), dim3(256), 0, c->stream, a);
  else if (comp) hipLaunchKernelGGL((k_source_idle<true, false>), dim3(gb), dim3(256), 0, c->stream, a);
  else if (fm) hipLaunchKernelGGL((k_source_idle<false, true>), dim3(gb), dim3(256), 0, c->stream, a);
  else hipLaunchKernelGGL((k_source_idle<false, false>), dim3(gb), dim3(256), 0, c->stream, a);
}

// A stage that also writes the adjoint tape (a.tape at the cursor's slot).
template <int ST>
void launch_stage_rec(inu_ctx* c, const Args& a, bool comp, bool fm) {
  const size_t lds = ST == 0 ? c->lds_stage0 : c->lds_stage1;
  const dim3 g(c->grid[ST][comp ? 1 : 0][fm ? 1 : 0][0]);
  launch_k_stage<ST, 0, true, false>(c, a, comp, fm, g, lds);
}

// Open edge words of stage ST (k_open_words), before that stage's kernel.
template <int ST>
void launch_open_words(inu_ctx* c, bool fm) {
  const dim3 g((c->n_open + 255) / 256), b(256);
  if (fm) hipLaunchKernelGGL((k_open_words<ST, true>), g, b, 0, c->stream, c->a, c->open_cell, c->open_geom);
  else hipLaunchKernelGGL((k_open_words<ST, false>), g, b, 0, c->stream, c->a, c->open_cell, c->open_geom);
}

int launch_step(inu_ctx* c, const hipEvent_t* ev) {
  const Args& a = c->a;
  const bool comp = (c->flags & INU_FLAG_COMPENSATED) != 0;
  const bool fm = (c->flags & INU_FLAG_EXACT_ORDER) == 0;
  if (c->n_open) launch_open_words<0>(c, fm);
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  launch_stage<0>(c, comp, fm);
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  if (c->n_open) launch_open_words<1>(c, fm);
  launch_stage<1>(c, comp, fm);
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  if (c->n_open)
    hipLaunchKernelGGL(k_open_tally, dim3((c->n_stretch + 63) / 64), dim3(64), 0, c->stream, a, c->stretch_ptr,
                       c->open_cell, c->tally, c->n_stretch);
  if (c->n_src) launch_source_idle(c, a, comp, fm);
  const uint32_t n_dil = (c->flags & INU_FLAG_ACTIVE_SET) ? a.n_dwg : 0u;
  const uint32_t n_fix = c->n_src ? (a.n_src_pairs + kFixChunk - 1) / kFixChunk : 0u;
  if (n_dil + n_fix) {
    if (fm) hipLaunchKernelGGL((k_step_tail<true>), dim3(n_dil + n_fix), dim3(256), 0, c->stream, a, n_dil);
    else hipLaunchKernelGGL((k_step_tail<false>), dim3(n_dil + n_fix), dim3(256), 0, c->stream, a, n_dil);
  }
  hipLaunchKernelGGL(k_step_control, dim3(1), dim3(256), 0, c->stream, a, 1);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Device address of `scratch` on each GPU, taken once per device with that
// device current (host code must use these, never the symbol's own name).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* ptr = nullptr;
  g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_ptr[device] = (g_scratch_err[device] == hipSuccess) ? ptr : nullptr;
  g_scratch_done[device] = true;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int inu_version(void) { return INU_ABI_VERSION; }

int inu_create(int device, inu_ctx** out) {
  if (!out) return INU_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return INU_ENODEV;
  if (device < 0 || device >= n) return INU_EINVAL;
  inu_ctx* c = new inu_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return INU_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  c->p = make_phys(9.81f, 1e-4f, 0.8f);
  scratch_setup(device);  // with the device current
  *out = c;
  return 0;
}

void inu_destroy(inu_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* inu_last_error(const inu_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"

namespace {

// One mesh as the caller hands it over (the arguments of inu_set_mesh).
struct MeshIn {
  uint32_t N, E;
  const int32_t *ecl, *ecr;
  const float *nx, *ny, *len, *dz, *area;
  const uint32_t *cptr, *cidx;
};

// Device layout of one mesh on its own (offsets and ids local to the mesh,
// BlockInfo.meta = real cells only), before it is placed in the union.
struct Layout {
  uint32_t N = 0, E = 0, S = 0, NB = 0, n_halo = 0;
  uint32_t max_local = 1, max_slots = 1, max_ghost = 1;
  std::vector<BlockInfo> binfo;
  std::vector<uint4> lrec;
  std::vector<float> ldz;
  std::vector<uint32_t> gx;
  std::vector<uint16_t> lidx;
  std::vector<int32_t> snb;
  std::vector<float> slen;
  std::vector<std::pair<uint32_t, uint32_t>> cross;
  std::vector<uint32_t> bptr, bidx;
  std::vector<uint32_t> stream_edge;  // local edge stream position -> edge
  std::vector<uint32_t> open_cell;    // open edge k -> its (inside) cell
  std::vector<float4> open_geom;      // open edge k -> (nx, ny, length, 0)
  std::vector<uint32_t> open_bptr;    // per block, its open edges in open_jk
  std::vector<uint2> open_jk;         // (local stream position, open edge index)
};

// Validates one mesh and builds its layout.
int build_layout(inu_ctx* c, const MeshIn& m, Layout& L) {
  if (m.N == 0 || m.E == 0 || !m.ecl || !m.ecr || !m.nx || !m.ny || !m.len || !m.dz || !m.area || !m.cptr || !m.cidx)
    return fail(c, INU_EINVAL, "empty mesh or null array");
  const uint32_t N = m.N, E = m.E, NB = (N + kBlockCells - 1) / kBlockCells;
  const int32_t* edge_cell_l = m.ecl;
  const int32_t* edge_cell_r = m.ecr;
  const float* edge_nx = m.nx;
  const float* edge_ny = m.ny;
  const float* edge_len = m.len;
  const float* edge_dz = m.dz;
  const uint32_t* cell_edge_ptr = m.cptr;
  const uint32_t* cell_edge_idx = m.cidx;
  const uint32_t S = cell_edge_ptr[N];
  if (cell_edge_ptr[0] != 0) return fail(c, INU_EINVAL, "cell_edge_ptr[0] must be 0");
  // Validate the edge order and the signed CSR.
  std::vector<uint32_t> eo(NB + 1, 0);
  for (uint32_t e = 0; e < E; ++e) {
    const int32_t l = edge_cell_l[e], r = edge_cell_r[e];
    if (l < 0 || (uint32_t)l >= N) return fail(c, INU_EINVAL, "edge_cell_l out of range");
    if (r >= (int64_t)N || (r < -1 && (int64_t)(-2 - (int64_t)r) >= (int64_t)E))
      return fail(c, INU_EINVAL, "edge_cell_r out of range (a cell, -1 for a wall, -2 - k for open edge k)");
    if (e > 0 && (uint32_t)l / kBlockCells < (uint32_t)edge_cell_l[e - 1] / kBlockCells)
      return fail(c, INU_EINVAL, "edges must be sorted by the block of their left cell");
    eo[(uint32_t)l / kBlockCells + 1]++;
  }
  for (uint32_t b = 0; b < NB; ++b) eo[b + 1] += eo[b];
  // Halo edges: right cell in b, left cell in another block; ascending ids.
  std::vector<uint32_t> hp(NB + 1, 0);
  for (uint32_t e = 0; e < E; ++e) {
    const int32_t r = edge_cell_r[e];
    if (r >= 0 && (uint32_t)r / kBlockCells != (uint32_t)edge_cell_l[e] / kBlockCells) hp[(uint32_t)r / kBlockCells + 1]++;
  }
  for (uint32_t b = 0; b < NB; ++b) hp[b + 1] += hp[b];
  std::vector<uint32_t> hx(hp[NB]);
  {
    std::vector<uint32_t> fill(hp.begin(), hp.end() - 1);
    for (uint32_t e = 0; e < E; ++e) {
      const int32_t r = edge_cell_r[e];
      if (r >= 0 && (uint32_t)r / kBlockCells != (uint32_t)edge_cell_l[e] / kBlockCells) hx[fill[(uint32_t)r / kBlockCells]++] = e;
    }
  }
  // Local edge stream of each block (its owned edges, then its halo edges)
  // with both sides as local cell numbers, and its ghost cells, ascending.
  for (uint32_t e = 0; e < E; ++e) {
    const float dzl = edge_dz[2 * e], dzr = edge_dz[2 * e + 1];
    if (!(dzl >= 0.0f) || !(dzr >= 0.0f) || !std::isfinite(dzl) || !std::isfinite(dzr) || (dzl != 0.0f && dzr != 0.0f))
      return fail(c, INU_EINVAL, "edge_dz: both steps must be finite and non-negative with one of them zero "
                                 "(the face sits on the higher bed)");
  }
  // Open edges: each index 0..n_open-1 once, on an edge without a bed step.
  {
    std::vector<int64_t> open_edge;
    for (uint32_t e = 0; e < E; ++e) {
      const int32_t r = edge_cell_r[e];
      if (r >= -1) continue;
      const uint32_t k = (uint32_t)(-2 - (int64_t)r);
      if (k >= open_edge.size()) open_edge.resize(k + 1, -1);
      if (open_edge[k] >= 0) return fail(c, INU_EINVAL, "two edges name the same open edge index");
      if (edge_dz[2 * e] != 0.0f || edge_dz[2 * e + 1] != 0.0f)
        return fail(c, INU_EINVAL, "an open boundary edge has a bed step");
      open_edge[k] = e;
    }
    L.open_cell.assign(open_edge.size(), 0u);
    L.open_geom.assign(open_edge.size(), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (size_t k = 0; k < open_edge.size(); ++k) {
      if (open_edge[k] < 0) return fail(c, INU_EINVAL, "open edge indices must be 0..n_open-1 with none missing");
      const int64_t e = open_edge[k];
      L.open_cell[k] = (uint32_t)edge_cell_l[e];
      L.open_geom[k] = make_float4(edge_nx[e], edge_ny[e], edge_len[e], 0.0f);
    }
  }
  uint32_t max_local = 1, max_slots = 1, max_ghost = 1;
  std::vector<BlockInfo>& binfo = L.binfo;
  binfo.assign(NB, BlockInfo{});
  std::vector<uint4>& lrec = L.lrec;
  lrec.assign((size_t)E + hp[NB], uint4{});
  std::vector<float>& ldz = L.ldz;
  ldz.assign((size_t)E + hp[NB], 0.0f);
  std::vector<uint32_t>& gxv = L.gx;
  gxv.clear();
  L.stream_edge.assign((size_t)E + hp[NB], 0u);
  std::vector<uint32_t> ghosts;
  L.open_bptr.assign(NB + 1, 0u);
  L.open_jk.clear();
  uint32_t e0 = 0;
  for (uint32_t b = 0; b < NB; ++b) {
    const uint32_t cb = b * kBlockCells, ce = std::min(cb + kBlockCells, N);
    const uint32_t n_own = eo[b + 1] - eo[b], n_loc = n_own + (hp[b + 1] - hp[b]);
    ghosts.clear();
    for (uint32_t e = eo[b]; e < eo[b + 1]; ++e)
      if (edge_cell_r[e] >= 0 && (uint32_t)edge_cell_r[e] / kBlockCells != b) ghosts.push_back((uint32_t)edge_cell_r[e]);
    for (uint32_t k = hp[b]; k < hp[b + 1]; ++k) ghosts.push_back((uint32_t)edge_cell_l[hx[k]]);
    std::sort(ghosts.begin(), ghosts.end());
    ghosts.erase(std::unique(ghosts.begin(), ghosts.end()), ghosts.end());
    if (ghosts.size() > kMaxGhosts) return fail(c, INU_ERANGE, "a block borders too many cells of other blocks");
    auto local = [&](int32_t cell) -> uint32_t {
      if ((uint32_t)cell / kBlockCells == b) return (uint32_t)cell - cb;
      return kBlockCells + (uint32_t)(std::lower_bound(ghosts.begin(), ghosts.end(), (uint32_t)cell) - ghosts.begin());
    };
    for (uint32_t j = 0; j < n_loc; ++j) {
      const uint32_t e = (j < n_own) ? eo[b] + j : hx[hp[b] + (j - n_own)];
      const uint32_t ql = local(edge_cell_l[e]);
      const uint32_t qr = (edge_cell_r[e] < 0) ? kSideWall : local(edge_cell_r[e]);
      if (edge_cell_r[e] < -1) L.open_jk.push_back(make_uint2(j, (uint32_t)(-2 - (int64_t)edge_cell_r[e])));
      uint32_t bx, by, bl;
      memcpy(&bx, &edge_nx[e], 4);
      memcpy(&by, &edge_ny[e], 4);
      memcpy(&bl, &edge_len[e], 4);
      lrec[e0 + j] = make_uint4((ql << 16) | qr, bx, by, bl);
      L.stream_edge[e0 + j] = e;
      ldz[e0 + j] = (edge_dz[2 * e] > 0.0f) ? edge_dz[2 * e] : -edge_dz[2 * e + 1];
    }
    binfo[b] = BlockInfo{e0, n_loc, n_own, (uint32_t)gxv.size(), (uint32_t)ghosts.size(), cell_edge_ptr[cb],
                         cell_edge_ptr[ce] - cell_edge_ptr[cb], ce - cb};
    L.open_bptr[b + 1] = (uint32_t)L.open_jk.size();
    gxv.insert(gxv.end(), ghosts.begin(), ghosts.end());
    e0 += n_loc;
    max_local = std::max(max_local, n_loc);
    max_slots = std::max(max_slots, binfo[b].n_slots);
    max_ghost = std::max(max_ghost, (uint32_t)ghosts.size());
  }
  if (max_local >= 32768u) return fail(c, INU_ERANGE, "a block touches more than 32767 edges");
  // Local CSR: each slot's edge as its position in its block's LDS edge list.
  std::vector<uint16_t>& lidx = L.lidx;
  lidx.assign(S, 0);
  std::vector<int32_t>& snb = L.snb;
  snb.assign(S, 0);
  std::vector<float>& slen = L.slen;
  slen.assign(S, 0.0f);
  std::vector<std::pair<uint32_t, uint32_t>>& cross = L.cross;
  cross.clear();
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t b = i / kBlockCells;
    if (cell_edge_ptr[i + 1] < cell_edge_ptr[i] || cell_edge_ptr[i + 1] > S)
      return fail(c, INU_EINVAL, "cell_edge_ptr must be non-decreasing and end at the slot count");
    for (uint32_t k = cell_edge_ptr[i]; k < cell_edge_ptr[i + 1]; ++k) {
      const uint32_t s = cell_edge_idx[k], e = s >> 1, right = s & 1u;
      if (e >= E) return fail(c, INU_EINVAL, "cell_edge_idx names an edge out of range");
      const int32_t own = right ? edge_cell_r[e] : edge_cell_l[e];
      if (own != (int32_t)i) return fail(c, INU_EINVAL, "cell_edge_idx entry does not belong to its cell");
      uint32_t j;
      if ((uint32_t)edge_cell_l[e] / kBlockCells == b) {
        j = e - eo[b];
      } else {
        const uint32_t* lo = hx.data() + hp[b];
        const uint32_t* hi = hx.data() + hp[b + 1];
        const uint32_t* f = std::lower_bound(lo, hi, e);
        if (f == hi || *f != e) return fail(c, INU_EINVAL, "internal: halo edge missing");
        j = (eo[b + 1] - eo[b]) + (uint32_t)(f - lo);
      }
      lidx[k] = (uint16_t)((j << 1) | right);
      snb[k] = right ? edge_cell_l[e] : edge_cell_r[e];
      slen[k] = edge_len[e];
      if (snb[k] >= 0 && (uint32_t)snb[k] / kBlockCells != b) cross.emplace_back(i, (uint32_t)snb[k]);
    }
  }
  // Two-ring block graph (reference blocks.py:63-118).
  std::vector<std::vector<uint32_t>> adj(NB);
  for (uint32_t e = 0; e < E; ++e) {
    const int32_t r = edge_cell_r[e];
    if (r < 0) continue;
    const uint32_t bl = (uint32_t)edge_cell_l[e] / kBlockCells, br = (uint32_t)r / kBlockCells;
    if (bl != br) {
      adj[bl].push_back(br);
      adj[br].push_back(bl);
    }
  }
  for (auto& v : adj) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  std::vector<uint32_t>& bptr = L.bptr;
  bptr.assign(NB + 1, 0);
  std::vector<uint32_t>& bidx = L.bidx;
  bidx.clear();
  for (uint32_t b = 0; b < NB; ++b) {
    std::vector<uint32_t> ring(adj[b]);
    for (uint32_t n : adj[b]) ring.insert(ring.end(), adj[n].begin(), adj[n].end());
    std::sort(ring.begin(), ring.end());
    ring.erase(std::unique(ring.begin(), ring.end()), ring.end());
    for (uint32_t n : ring)
      if (n != b) bidx.push_back(n);
    bptr[b + 1] = (uint32_t)bidx.size();
  }
  L.N = N;
  L.E = E;
  L.S = S;
  L.NB = NB;
  L.n_halo = hp[NB];
  L.max_local = max_local;
  L.max_slots = max_slots;
  L.max_ghost = max_ghost;
  return 0;
}

// Places scenario s's copy of mesh scen_mesh[s] at block offset b0[s] of the
// union and uploads the union. Scenario s owns caller cells
// [cell0[s], cell0[s + 1]) and internal cells [256 b0[s], 256 b0[s] + N).
int assemble(inu_ctx* c, const std::vector<MeshIn>& ins, std::vector<Layout>& lay, const std::vector<uint32_t>& scen_mesh) {
  const uint32_t n_scen = (uint32_t)scen_mesh.size();
  if (n_scen == 0) return fail(c, INU_EINVAL, "no scenarios");
  if (n_scen > (1u << 22)) return fail(c, INU_ERANGE, "too many scenarios");
  std::vector<uint32_t> cell0(n_scen + 1, 0), block0(n_scen + 1, 0);
  uint64_t nb_tot = 0, e_tot = 0, s_tot = 0, l_tot = 0, g_tot = 0, cross_tot = 0, ring_tot = 0, n_tot = 0;
  uint32_t max_local = 1, max_slots = 1, max_ghost = 1, n_halo = 0;
  for (uint32_t s = 0; s < n_scen; ++s) {
    if (scen_mesh[s] >= lay.size()) return fail(c, INU_EINVAL, "scenario_mesh names a mesh out of range");
    const Layout& L = lay[scen_mesh[s]];
    cell0[s] = (uint32_t)n_tot;
    block0[s] = (uint32_t)nb_tot;
    n_tot += L.N;
    nb_tot += L.NB;
    e_tot += L.E;
    s_tot += L.S;
    l_tot += L.lrec.size();
    g_tot += L.gx.size();
    cross_tot += L.cross.size();
    ring_tot += L.bidx.size();
    n_halo += L.n_halo;
    max_local = std::max(max_local, L.max_local);
    max_slots = std::max(max_slots, L.max_slots);
    max_ghost = std::max(max_ghost, L.max_ghost);
  }
  // The edge words' LDS region (16 bytes per edge in the predictor) also
  // holds the unordered sums of a block's cells (one float4 each).
  max_local = std::max(max_local, kBlockCells);
  cell0[n_scen] = (uint32_t)n_tot;
  block0[n_scen] = (uint32_t)nb_tot;
  uint32_t n_open = 0;
  for (uint32_t s = 0; s < n_scen; ++s) n_open += (uint32_t)lay[scen_mesh[s]].open_cell.size();
  if (n_open && n_scen != 1) return fail(c, INU_EINVAL, "open boundaries take a context of one scenario");
  // Cell ids are int32 and every device array stays below 4 GiB.
  if (nb_tot * kBlockCells > 0x7fffffffull || l_tot * sizeof(uint4) >= (1ull << 32) || s_tot >= (1ull << 30) ||
      e_tot > 0xffffffffull)
    return fail(c, INU_ERANGE, "the meshes are too large for one context");
  const uint32_t NB = (uint32_t)nb_tot;
  const size_t NP = (size_t)NB * kBlockCells;
  std::vector<BlockInfo> binfo(NB);
  std::vector<uint4> lrec;
  lrec.reserve(l_tot);
  std::vector<float> ldz;
  ldz.reserve(l_tot);
  std::vector<uint32_t> gx;
  gx.reserve(g_tot);
  std::vector<uint32_t> cptr(NP + 1, 0);
  std::vector<uint16_t> lidx;
  lidx.reserve(s_tot);
  std::vector<int32_t> snb;
  snb.reserve(s_tot);
  std::vector<float> slen;
  slen.reserve(s_tot);
  std::vector<std::pair<uint32_t, uint32_t>> cross;
  cross.reserve(cross_tot);
  std::vector<uint32_t> bptr(NB + 1, 0), bidx;
  bidx.reserve(ring_tot);
  std::vector<float> area(NP, 0.0f);
  // The fast order's inverse areas: 1 / A of every real cell (the same float
  // division as the oracle's), 0 in the padding.
  std::vector<float> inv_area(NP, 0.0f);
  std::vector<int32_t> e_cl, e_cr;
  e_cl.reserve(e_tot);
  e_cr.reserve(e_tot);
  std::vector<uint32_t> stream_edge;
  stream_edge.reserve(l_tot);
  std::vector<uint32_t> own_pos(e_tot, 0u);
  uint32_t s_off = 0, ed_off = 0;
  for (uint32_t s = 0; s < n_scen; ++s) {
    const Layout& L = lay[scen_mesh[s]];
    const MeshIn& m = ins[scen_mesh[s]];
    const uint32_t b0 = block0[s], c0 = b0 * kBlockCells;
    const uint32_t e_off = (uint32_t)lrec.size(), g_off = (uint32_t)gx.size();
    for (uint32_t b = 0; b < L.NB; ++b) {
      BlockInfo bi = L.binfo[b];
      bi.e0 += e_off;
      bi.g0 += g_off;
      bi.s0 += s_off;
      bi.meta = (s << 9) | bi.meta;
      binfo[b0 + b] = bi;
      bptr[b0 + b + 1] = bptr[b0 + b] + (L.bptr[b + 1] - L.bptr[b]);
    }
    for (uint32_t n : L.bidx) bidx.push_back(b0 + n);
    for (uint32_t b = 0; b < L.NB; ++b)
      for (uint32_t j = 0; j < L.binfo[b].n_own; ++j)
        own_pos[ed_off + L.stream_edge[L.binfo[b].e0 + j]] = e_off + L.binfo[b].e0 + j;
    for (uint32_t e : L.stream_edge) stream_edge.push_back(ed_off + e);
    for (uint32_t e = 0; e < L.E; ++e) {
      e_cl.push_back((int32_t)(c0 + (uint32_t)m.ecl[e]));
      e_cr.push_back(m.ecr[e] < 0 ? m.ecr[e] : (int32_t)(c0 + (uint32_t)m.ecr[e]));
    }
    ed_off += L.E;
    lrec.insert(lrec.end(), L.lrec.begin(), L.lrec.end());
    ldz.insert(ldz.end(), L.ldz.begin(), L.ldz.end());
    for (uint32_t g : L.gx) gx.push_back(c0 + g);
    for (uint32_t i = 0; i < L.NB * kBlockCells; ++i) cptr[c0 + i + 1] = s_off + m.cptr[std::min(i + 1, L.N)];
    lidx.insert(lidx.end(), L.lidx.begin(), L.lidx.end());
    for (int32_t n : L.snb) snb.push_back(n < 0 ? n : (int32_t)(c0 + (uint32_t)n));
    slen.insert(slen.end(), L.slen.begin(), L.slen.end());
    for (const auto& pr : L.cross) cross.emplace_back(c0 + pr.first, c0 + pr.second);
    std::copy(m.area, m.area + L.N, area.begin() + c0);
    for (uint32_t i = 0; i < L.N; ++i) inv_area[c0 + i] = 1.0f / m.area[i];
    s_off += L.S;
  }

  free_all(c);
  c->N = (uint32_t)n_tot;
  c->E = (uint32_t)e_tot;
  c->S = (uint32_t)s_tot;
  c->NB = NB;
  c->max_local = max_local;
  c->n_halo = n_halo;
  c->n_scen = n_scen;
  c->cell0 = cell0;
  c->block0 = block0;
  c->n_src = 0;
  c->n_src_blocks = 0;
  c->n_src_pairs = 0;
  Args& a = c->a;
  int rc = 0;
  rc |= dalloc(c, &a.h, NP); rc |= dalloc(c, &a.hu, NP); rc |= dalloc(c, &a.hv, NP); rc |= dalloc(c, &a.hlo, NP);
  rc |= dalloc(c, &a.h1, NP); rc |= dalloc(c, &a.hu1, NP); rc |= dalloc(c, &a.hv1, NP); rc |= dalloc(c, &a.h1lo, NP);
  rc |= dalloc(c, &a.area, NP); rc |= dalloc(c, &a.nman, NP); rc |= dalloc(c, &a.clamp, NP); rc |= dalloc(c, &a.src, NP);
  rc |= dalloc(c, &a.inv_area, NP);
  uint32_t* dcptr; uint16_t* dlid; int32_t* dsn; float* dsl;
  rc |= dalloc(c, &dcptr, NP + 1); rc |= dalloc(c, &dlid, lidx.size()); rc |= dalloc(c, &dsn, snb.size());
  rc |= dalloc(c, &dsl, slen.size());
  uint4* dlrec; float* dldz;
  rc |= dalloc(c, &dlrec, lrec.size()); rc |= dalloc(c, &dldz, ldz.size());
  uint32_t *dgx, *dbptr, *dbidx;
  BlockInfo* dbinfo;
  rc |= dalloc(c, &dbinfo, NB); rc |= dalloc(c, &dgx, gx.size());
  rc |= dalloc(c, &dbptr, NB + 1); rc |= dalloc(c, &dbidx, bidx.size());
  rc |= dalloc(c, &a.wet, NB); rc |= dalloc(c, &a.active, NB); rc |= dalloc(c, &a.forced, NB);
  rc |= dalloc(c, &a.has_src, NB); rc |= dalloc(c, &a.list, NB); rc |= dalloc(c, &a.idle, NB);
  rc |= dalloc(c, &a.part_dry, NB);
  a.n_dwg = (NB + kDilateBlocks - 1) / kDilateBlocks;
  rc |= dalloc(c, &a.act_part, 3 * (size_t)a.n_dwg);
  uint32_t* src_cell; uint32_t* src_blocks;
  uint2* src_pair;
  uint32_t* fix_blk;
  uint2* fix_chunk;
  rc |= dalloc(c, &src_cell, 1); rc |= dalloc(c, &src_blocks, 1); rc |= dalloc(c, &src_pair, 1);
  rc |= dalloc(c, &fix_blk, 1); rc |= dalloc(c, &fix_chunk, 1);
  rc |= dalloc(c, &c->ctl, 1);
  rc |= dalloc(c, &a.walk, 4 * kXcds * kWalkStride);
  rc |= dalloc(c, &c->scen, n_scen);
  c->n_open = n_open;
  c->n_stretch = n_open ? 1u : 0u;
  c->open_cell_h.clear();
  c->open_cell = c->stretch_ptr = nullptr;
  c->tally = nullptr;
  c->open_geom = nullptr;
  c->open_h = OpenDev{};
  a.open = nullptr;
  a.n_open = n_open;
  if (n_open) {
    OpenDev& o = c->open_h;
    uint32_t *okind, *obptr;
    float *oval, *oflux;
    float4* owords;
    uint2* ojk;
    OpenDev* odev;
    rc |= dalloc(c, &okind, n_open); rc |= dalloc(c, &oval, n_open); rc |= dalloc(c, &oflux, 2 * (size_t)n_open);
    rc |= dalloc(c, &owords, 4 * (size_t)n_open); rc |= dalloc(c, &obptr, NB + 1); rc |= dalloc(c, &ojk, n_open);
    rc |= dalloc(c, &odev, 1);
    rc |= dalloc(c, &c->open_cell, n_open); rc |= dalloc(c, &c->stretch_ptr, n_open + 1);
    rc |= dalloc(c, &c->tally, 4 * (size_t)n_open); rc |= dalloc(c, &c->open_geom, n_open);
    o = OpenDev{okind, oval, owords, oflux, obptr, ojk};
    a.open = odev;
    for (uint32_t cl : lay[scen_mesh[0]].open_cell) c->open_cell_h.push_back(cl);  // one scenario: no offset
  }
  if (rc) return rc;
  a.cptr = dcptr; a.lidx = dlid; a.snb = dsn; a.slen = dsl;
  a.lrec = dlrec; a.ldz = dldz;
  a.binfo = dbinfo; a.gx = dgx; a.bptr = dbptr; a.bidx = dbidx;
  a.src_cell = src_cell; a.src_blocks = src_blocks; a.src_pair = src_pair; a.n_src_pairs = 0;
  a.fix_blk = fix_blk; a.fix_chunk = fix_chunk; a.n_fix = 0; a.src_neg = 0;
  a.ctl = c->ctl;
  a.scen = c->scen;
  a.n_scen = n_scen;
  a.N = (uint32_t)NP; a.NB = NB; a.max_local = max_local; a.max_slots = max_slots; a.max_ghost = max_ghost;
  a.n_lrec = (uint32_t)std::max<size_t>(lrec.size(), 1); a.n_lidx = (uint32_t)std::max<size_t>(lidx.size(), 1);
  HIPCHK(c, hipMemsetAsync(a.h, 0, NP * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.hu, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.hv, 0, NP * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.hlo, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.h1, 0, NP * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.hu1, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.hv1, 0, NP * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.h1lo, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.clamp, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.src, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.wet, 0, NB * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.active, 0, NB * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.forced, 0, NB * 4, c->stream)); HIPCHK(c, hipMemsetAsync(a.has_src, 0, NB * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.part_dry, 1, NB * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(a.act_part, 0, 3 * (size_t)a.n_dwg * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->ctl, 0, sizeof(Ctl), c->stream));
  HIPCHK(c, hipMemsetAsync(a.walk, 0, 4 * kXcds * kWalkStride * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(c->scen, 0, sizeof(Scen) * n_scen, c->stream));
  if (n_open) {
    // Every open edge starts closed (a wall) and the mesh is one stretch.
    const OpenDev& o = c->open_h;
    if ((rc = upload(c, const_cast<OpenDev*>(a.open), &o, 1))) return rc;
    HIPCHK(c, hipMemsetAsync(const_cast<uint32_t*>(o.kind), 0, n_open * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(const_cast<float*>(o.value), 0, n_open * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(o.flux, 0, 2 * (size_t)n_open * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->tally, 0, 4 * (size_t)n_open * 4, c->stream));
    const uint32_t ptr2[2] = {0u, n_open};
    if ((rc = upload(c, c->stretch_ptr, ptr2, 2))) return rc;
    if ((rc = upload(c, c->open_cell, c->open_cell_h.data(), n_open))) return rc;
    if ((rc = upload(c, c->open_geom, lay[scen_mesh[0]].open_geom.data(), n_open))) return rc;
    if ((rc = upload(c, const_cast<uint32_t*>(o.bptr), lay[scen_mesh[0]].open_bptr.data(), NB + 1))) return rc;
    if ((rc = upload(c, const_cast<uint2*>(o.jk), lay[scen_mesh[0]].open_jk.data(), n_open))) return rc;
    HIPCHK(c, hipMemsetAsync(o.words, 0, 4 * (size_t)n_open * sizeof(float4), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  std::vector<float> nm(NP, 0.03f);
  if ((rc = upload(c, a.nman, nm.data(), NP))) return rc;
  if ((rc = upload(c, a.area, area.data(), NP))) return rc;
  if ((rc = upload(c, a.inv_area, inv_area.data(), NP))) return rc;
  if ((rc = upload(c, dcptr, cptr.data(), NP + 1))) return rc;
  {
    std::vector<uint16_t> lend(NP, 0);
    for (size_t i = 0; i < NP; ++i) lend[i] = (uint16_t)(cptr[i + 1] - binfo[i / kBlockCells].s0);
    uint16_t* dlend;
    if ((rc = dalloc(c, &dlend, NP))) return rc;
    if ((rc = upload(c, dlend, lend.data(), NP))) return rc;
    a.lend = dlend;
  }
  if ((rc = upload(c, dlid, lidx.data(), lidx.size()))) return rc;
  if ((rc = upload(c, dsn, snb.data(), snb.size()))) return rc;
  if ((rc = upload(c, dsl, slen.data(), slen.size()))) return rc;
  if ((rc = upload(c, dlrec, lrec.data(), lrec.size()))) return rc;
  if ((rc = upload(c, dldz, ldz.data(), ldz.size()))) return rc;
  if ((rc = upload(c, dbinfo, binfo.data(), NB))) return rc;
  if ((rc = upload(c, dgx, gx.data(), gx.size()))) return rc;
  if ((rc = upload(c, dbptr, bptr.data(), NB + 1))) return rc;
  if ((rc = upload(c, dbidx, bidx.data(), bidx.size()))) return rc;
  // Grouped by the neighbour's block once here, so that every source list
  // keeps the groups the source fix walks without sorting again.
  std::stable_sort(cross.begin(), cross.end(), [](const std::pair<uint32_t, uint32_t>& x,
                                                  const std::pair<uint32_t, uint32_t>& y) {
    return x.second / kBlockCells < y.second / kBlockCells;
  });
  c->cross.swap(cross);
  c->e_cl.swap(e_cl);
  c->e_cr.swap(e_cr);
  c->stream_edge.swap(stream_edge);
  c->own_pos.swap(own_pos);
  c->h_ldz.swap(ldz);
  c->src_nonneg = true;
  c->lds_stage0 = lds_words_stage0(max_local, max_slots, max_ghost) * sizeof(float);
  c->lds_stage1 = lds_words_stage1(max_local, max_slots, max_ghost) * sizeof(float);
  c->lds_full = lds_words(max_local, max_slots, max_ghost) * sizeof(float);
  if (c->lds_full > 160 * 1024) return fail(c, INU_ERANGE, "a block's edge list does not fit in LDS");
  // Persistent grids: as many workgroups as fit at once, a multiple of the
  // XCD count, per stage kernel variant.
  const void* variants[2][2][2][kTiles] = {
      {{{(const void*)k_stage<0, false, false, 0>, (const void*)k_stage<0, false, false, 1>,
         (const void*)k_stage<0, false, false, 2>},
        {(const void*)k_stage<0, false, true, 0>, (const void*)k_stage<0, false, true, 1>,
         (const void*)k_stage<0, false, true, 2>}},
       {{(const void*)k_stage<0, true, false, 0>, (const void*)k_stage<0, true, false, 1>,
         (const void*)k_stage<0, true, false, 2>},
        {(const void*)k_stage<0, true, true, 0>, (const void*)k_stage<0, true, true, 1>,
         (const void*)k_stage<0, true, true, 2>}}},
      {{{(const void*)k_stage<1, false, false, 0>, (const void*)k_stage<1, false, false, 1>,
         (const void*)k_stage<1, false, false, 2>},
        {(const void*)k_stage<1, false, true, 0>, (const void*)k_stage<1, false, true, 1>,
         (const void*)k_stage<1, false, true, 2>}},
       {{(const void*)k_stage<1, true, false, 0>, (const void*)k_stage<1, true, false, 1>,
         (const void*)k_stage<1, true, false, 2>},
        {(const void*)k_stage<1, true, true, 0>, (const void*)k_stage<1, true, true, 1>,
         (const void*)k_stage<1, true, true, 2>}}}};
  for (int st = 0; st < 2; ++st)
    for (int cp = 0; cp < 2; ++cp)
      for (int fm = 0; fm < 2; ++fm)
        for (int tl = 0; tl < kTiles; ++tl) {
          int occ = 0;
          HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, variants[st][cp][fm][tl], 256,
                                                                  st ? c->lds_stage1 : c->lds_stage0));
          // Tile 0: no more workgroups per CU than the kernel is compiled
          // for (a seventh predictor workgroup fits in its registers and
          // LDS but adds memory waits without adding throughput).
          if (tl == 0) occ = std::min(occ, st ? stage_wgs<1>() : stage_wgs<0>());
          uint32_t g = (uint32_t)(std::max(occ, 1) * c->cus);
          g = std::max<uint32_t>(kXcds, std::min<uint32_t>(g, ((NB + kXcds - 1) / kXcds) * kXcds));
          c->grid[st][cp][fm][tl] = (g / kXcds) * kXcds;
        }
  c->tile = 0;
  c->grid_stage = c->grid[1][1][1][0];
  {
    // The prologue keeps the two-record cell layout (carve()); its grid is
    // what fits with that LDS.
    int occ = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_prologue_cfl<true, true>, 256,
                                                            c->lds_full));
    c->grid_prologue = std::max<uint32_t>(1u, (uint32_t)std::max(occ, 1) * c->cus);
  }
  c->a.p = c->p;
  c->has_mesh = true;
  return 0;
}

}  // namespace

extern "C" {

int inu_set_mesh(inu_ctx* c, uint32_t n_cells, uint32_t n_edges, const int32_t* edge_cell_l,
                 const int32_t* edge_cell_r, const float* edge_nx, const float* edge_ny,
                 const float* edge_len, const float* edge_dz, const float* cell_area,
                 const uint32_t* cell_edge_ptr, const uint32_t* cell_edge_idx) {
  if (!c) return INU_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<MeshIn> ins{MeshIn{n_cells, n_edges, edge_cell_l, edge_cell_r, edge_nx, edge_ny, edge_len, edge_dz,
                                 cell_area, cell_edge_ptr, cell_edge_idx}};
  std::vector<Layout> lay(1);
  int rc = build_layout(c, ins[0], lay[0]);
  if (rc) return rc;
  return assemble(c, ins, lay, std::vector<uint32_t>{0u});
}

int inu_set_batch(inu_ctx* c, uint32_t n_meshes, const inu_mesh* meshes, uint32_t n_scenarios,
                  const uint32_t* scenario_mesh) {
  if (!c) return INU_EINVAL;
  if (n_meshes == 0 || !meshes || n_scenarios == 0 || !scenario_mesh)
    return fail(c, INU_EINVAL, "inu_set_batch: no meshes or no scenarios");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<MeshIn> ins(n_meshes);
  std::vector<Layout> lay(n_meshes);
  for (uint32_t k = 0; k < n_meshes; ++k) {
    const inu_mesh& m = meshes[k];
    ins[k] = MeshIn{m.n_cells, m.n_edges, m.edge_cell_l, m.edge_cell_r, m.nx, m.ny, m.len, m.edge_dz,
                    m.cell_area, m.cell_edge_ptr, m.cell_edge_idx};
    int rc = build_layout(c, ins[k], lay[k]);
    if (rc) {
      c->err = "mesh " + std::to_string(k) + ": " + c->err;
      return rc;
    }
  }
  return assemble(c, ins, lay, std::vector<uint32_t>(scenario_mesh, scenario_mesh + n_scenarios));
}

int inu_scenario_count(inu_ctx* c) {
  int rc = check_ready(c);
  if (rc) return rc;
  return (int)c->n_scen;
}

int inu_scenario_info(inu_ctx* c, uint32_t s, uint32_t* out4) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out4) return fail(c, INU_EINVAL, "null output");
  if (s >= c->n_scen) return fail(c, INU_EINVAL, "scenario out of range");
  out4[0] = c->cell0[s];
  out4[1] = c->cell0[s + 1] - c->cell0[s];
  out4[2] = c->block0[s];
  out4[3] = c->block0[s + 1] - c->block0[s];
  return 0;
}

int inu_mesh_info(inu_ctx* c, uint32_t* out6) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out6) return fail(c, INU_EINVAL, "null output");
  out6[0] = c->N; out6[1] = c->E; out6[2] = c->S; out6[3] = c->NB; out6[4] = c->max_local;
  out6[5] = c->grid_stage;
  return 0;
}

int inu_set_params(inu_ctx* c, float g, float dry_tol, float cfl, uint32_t flags) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!(g > 0.0f) || !(dry_tol > 0.0f) || !(cfl > 0.0f) || !std::isfinite(g) || !std::isfinite(dry_tol) || !std::isfinite(cfl))
    return fail(c, INU_EINVAL, "g, dry_tol and cfl must be finite and positive");
  if (!(flags & INU_FLAG_DETERMINISTIC) && c->n_open)
    return fail(c, INU_EINVAL, "open boundaries need deterministic accumulation");
  // The fast order's reciprocals take thin-film divisors h^2 + (4 dry_tol)^2,
  // exact while they lie in [2^-126, 2^126].
  if (!(flags & INU_FLAG_EXACT_ORDER) && !(dry_tol >= 1e-18f && dry_tol <= 1e18f))
    return fail(c, INU_EINVAL, "dry_tol must lie in [1e-18, 1e18] (the exact order takes any positive value)");
  if ((flags ^ c->flags) & ~INU_FLAG_TRACK_CLAMP) {
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
  }
  c->p = make_phys(g, dry_tol, cfl);
  c->flags = flags;
  c->a.p = c->p;
  c->a.track_clamp = (flags & INU_FLAG_TRACK_CLAMP) ? 1u : 0u;
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  uint32_t swap = (flags & INU_FLAG_ACTIVE_SET) ? 1u : 0u;
  HIPCHK(c, hipMemcpyAsync(&c->ctl->swap_on, &swap, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_set_state(inu_ctx* c, const float* h, const float* hu, const float* hv, const float* h_lo) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!h || !hu || !hv) return fail(c, INU_EINVAL, "inu_set_state: null array");
  if ((rc = upload_cells(c, c->a.h, h))) return rc;
  if ((rc = depths_replaced(c))) return rc;
  if ((rc = upload_cells(c, c->a.hu, hu))) return rc;
  if ((rc = upload_cells(c, c->a.hv, hv))) return rc;
  if (h_lo) {
    if ((rc = upload_cells(c, c->a.hlo, h_lo))) return rc;
  } else {
    HIPCHK(c, hipMemsetAsync(c->a.hlo, 0, (size_t)c->NB * kBlockCells * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  // A new initial state starts the open-boundary volume tally again
  // (reference gpu_resident_batch.py:440-445).
  if (c->n_open) {
    HIPCHK(c, hipMemsetAsync(c->tally, 0, 4 * (size_t)c->n_stretch * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int inu_set_manning(inu_ctx* c, const float* n) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!n) return fail(c, INU_EINVAL, "inu_set_manning: null array");
  for (uint32_t i = 0; i < c->N; ++i)
    if (!std::isfinite(n[i]) || n[i] < 0.0f) return fail(c, INU_EINVAL, "manning n must be finite and non-negative");
  return upload_cells(c, c->a.nman, n);
}

int inu_set_sources(inu_ctx* c, uint32_t n, const uint32_t* cell, const float* rate) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n && (!cell || !rate)) return fail(c, INU_EINVAL, "inu_set_sources: null array");
  if (c->adj.on) return fail(c, INU_ESTATE, "inu_set_sources during an adjoint recording (use inu_set_source_rates)");
  std::vector<float> dense((size_t)c->NB * kBlockCells, 0.0f);
  std::vector<uint32_t> has(c->NB, 0), cells;
  cells.reserve(n);
  for (uint32_t k = 0; k < n; ++k) {
    if (cell[k] >= c->N) return fail(c, INU_EINVAL, "source cell out of range");
    if (!std::isfinite(rate[k])) return fail(c, INU_EINVAL, "source rate must be finite");
    if (rate[k] == 0.0f) continue;
    const uint32_t ic = internal_cell(c, cell[k]);
    if (dense[ic] != 0.0f) return fail(c, INU_EINVAL, "a source cell is listed twice");
    dense[ic] = rate[k];
    cells.push_back(ic);
    has[ic / kBlockCells] = 1;
  }
  std::sort(cells.begin(), cells.end());
  std::vector<uint2> pairs;
  for (const auto& pr : c->cross)
    if (dense[pr.first] != 0.0f) pairs.push_back(make_uint2(pr.first, pr.second));
  // The groups of pairs whose neighbours share a block (c->cross is in that
  // order), and the groups each chunk of kFixChunk pairs touches, for
  // the source fix.
  std::vector<uint32_t> fix_blk, pair_group(pairs.size());
  for (uint32_t k = 0; k < pairs.size(); ++k) {
    if (k == 0 || pairs[k].y / kBlockCells != pairs[k - 1].y / kBlockCells) fix_blk.push_back(pairs[k].y / kBlockCells);
    pair_group[k] = (uint32_t)fix_blk.size() - 1u;
  }
  std::vector<uint2> fix_chunk;
  for (size_t k = 0; k < pairs.size(); k += kFixChunk) {
    const size_t last = std::min(pairs.size(), k + kFixChunk) - 1;
    fix_chunk.push_back(make_uint2(pair_group[k], pair_group[last] + 1u));
  }
  std::vector<uint32_t> blocks;
  for (uint32_t b = 0; b < c->NB; ++b)
    if (has[b]) blocks.push_back(b);
  // The new lists go to fresh device buffers first; the old ones stay in
  // place (with their counts) until every allocation and upload of the new
  // ones has succeeded, so a failure leaves the context as it was.
  struct List {
    const void* host;
    size_t bytes;
    void* dev;
  };
  List fresh[5] = {{cells.data(), cells.size() * 4, nullptr},
                   {blocks.data(), blocks.size() * 4, nullptr},
                   {pairs.data(), pairs.size() * 8, nullptr},
                   {fix_blk.data(), fix_blk.size() * 4, nullptr},
                   {fix_chunk.data(), fix_chunk.size() * 8, nullptr}};
  hipError_t e = hipSuccess;
  for (List& f : fresh) {
    if (e == hipSuccess) e = hipMalloc(&f.dev, f.bytes ? f.bytes : 16);
    if (e == hipSuccess && f.bytes) e = hipMemcpyAsync(f.dev, f.host, f.bytes, hipMemcpyHostToDevice, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    for (List& f : fresh)
      if (f.dev) (void)hipFree(f.dev);
    return fail(c, INU_EHIP, std::string("inu_set_sources: ") + hipGetErrorString(e));
  }
  // Swap: free the old lists (the stream is idle after the synchronisation).
  for (void* q : {(void*)c->a.src_cell, (void*)c->a.src_blocks, (void*)c->a.src_pair, (void*)c->a.fix_blk,
                  (void*)c->a.fix_chunk}) {
    auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
    if (it != c->allocs.end()) {
      (void)hipFree(q);
      c->allocs.erase(it);
    }
  }
  for (List& f : fresh) c->allocs.push_back(f.dev);
  c->a.src_cell = static_cast<const uint32_t*>(fresh[0].dev);
  c->a.src_blocks = static_cast<const uint32_t*>(fresh[1].dev);
  c->a.src_pair = static_cast<const uint2*>(fresh[2].dev);
  c->a.fix_blk = static_cast<const uint32_t*>(fresh[3].dev);
  c->a.fix_chunk = static_cast<const uint2*>(fresh[4].dev);
  c->a.n_fix = (uint32_t)fix_blk.size();
  c->n_src_pairs = (uint32_t)pairs.size();
  c->a.n_src_pairs = c->n_src_pairs;
  c->n_src = (uint32_t)cells.size();
  c->n_src_blocks = (uint32_t)blocks.size();
  c->a.n_src = c->n_src;
  c->a.n_src_blocks = c->n_src_blocks;
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  // The dense rates and per-block flags live in buffers of fixed size; if
  // their upload fails, the context runs with no sources rather than with
  // lists that disagree with the rates.
  rc = upload(c, c->a.src, dense.data(), dense.size());
  if (!rc) rc = upload(c, c->a.has_src, has.data(), has.size());
  if (rc) {
    c->n_src = c->n_src_blocks = c->n_src_pairs = 0;
    c->a.n_src = c->a.n_src_blocks = c->a.n_src_pairs = c->a.n_fix = 0;
    return rc;
  }
  c->src_cells = cells;
  c->src_nonneg = true;
  for (uint32_t k = 0; k < n; ++k)
    if (rate[k] < 0.0f) c->src_nonneg = false;
  c->a.src_neg = c->src_nonneg ? 0u : 1u;
  return 0;
}

int inu_ctrl_write(inu_ctx* c, const inu_ctrl* ctrl) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!ctrl) return fail(c, INU_EINVAL, "null ctrl");
  std::vector<Scen> sc;
  if ((rc = scen_read(c, sc))) return rc;
  for (uint32_t s = 0; s < c->n_scen; ++s) sc[s].ctrl = ctrl[s];
  return copy_sync(c, c->scen, sc.data(), sizeof(Scen) * c->n_scen, hipMemcpyHostToDevice);
}

int inu_ctrl_read(inu_ctx* c, inu_ctrl* ctrl) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!ctrl) return fail(c, INU_EINVAL, "null ctrl");
  std::vector<Scen> sc;
  if ((rc = scen_read(c, sc))) return rc;
  for (uint32_t s = 0; s < c->n_scen; ++s) ctrl[s] = sc[s].ctrl;
  return 0;
}

int inu_dt_words(inu_ctx* c, uint32_t* out4) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out4) return fail(c, INU_EINVAL, "null output");
  std::vector<Scen> sc;
  if ((rc = scen_read(c, sc))) return rc;
  for (uint32_t s = 0; s < c->n_scen; ++s) memcpy(out4 + 4 * s, &sc[s].dt_cur, 16);
  return 0;
}

int inu_activity(inu_ctx* c, uint64_t* out4, int reset) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out4) return fail(c, INU_EINVAL, "null output");
  Ctl h;
  HIPCHK(c, hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; ++k) out4[k] = h.act_total[k];
  out4[3] = h.act_steps;
  if (reset) {
    unsigned long long z[4] = {0ull, 0ull, 0ull, 0ull};
    HIPCHK(c, hipMemcpyAsync(&c->ctl->act_total[0], z, sizeof(z), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int inu_prologue(inu_ctx* c) {
  int rc = check_ready(c);
  if (rc) return rc;
  const Args& a = c->a;
  const bool act = (c->flags & INU_FLAG_ACTIVE_SET) != 0;
  const uint32_t gb = (c->NB + 255) / 256;
  const uint32_t gm = (std::max(c->NB, c->n_scen) + 255) / 256;
  hipLaunchKernelGGL(k_block_mark, dim3(gm), dim3(256), 0, c->stream, a, act ? 0 : 1);
  hipLaunchKernelGGL(k_dilate, dim3(a.n_dwg), dim3(256), 0, c->stream, a, 1, act ? 0 : 1);
  const uint32_t gp = std::min<uint32_t>(c->NB, c->grid_prologue);
  if (c->flags & INU_FLAG_EXACT_ORDER)
    hipLaunchKernelGGL((k_prologue_cfl<false, false>), dim3(gp), dim3(256), c->lds_full, c->stream, a);
  else if (c->flags & INU_FLAG_COMPENSATED)
    hipLaunchKernelGGL((k_prologue_cfl<true, true>), dim3(gp), dim3(256), c->lds_full, c->stream, a);
  else
    hipLaunchKernelGGL((k_prologue_cfl<true, false>), dim3(gp), dim3(256), c->lds_full, c->stream, a);
  hipLaunchKernelGGL(k_step_control, dim3(1), dim3(256), 0, c->stream, a, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_retarget(inu_ctx* c) {
  int rc = check_ready(c);
  if (rc) return rc;
  // The lists again hold every block whose flag is set (scenarios at their
  // stop dropped theirs); the flags themselves are kept.
  const uint32_t gb = (c->NB + 255) / 256;
  hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(1), 0, c->stream, c->a);
  hipLaunchKernelGGL(k_dilate, dim3(c->a.n_dwg), dim3(256), 0, c->stream, c->a, 1, 1);
  hipLaunchKernelGGL(k_step_control, dim3(1), dim3(256), 0, c->stream, c->a, 2);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_steps(inu_ctx* c, uint32_t n, int use_graph) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n == 0) return 0;
  if (!use_graph) {
    for (uint32_t k = 0; k < n; ++k)
      if ((rc = launch_step(c, nullptr))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  // Runs of more than kGraphSteps steps replay the kGraphSteps-step graph
  // (and one of the remainder), launched back to back with one wait at the
  // end: graphs stay small whatever n is.
  for (uint32_t left = n; left != 0u;) {
    const uint32_t m = std::min(left, kGraphSteps);
    auto it = c->graphs.find(m);
    if (it == c->graphs.end()) {
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (uint32_t k = 0; k < m; ++k) {
        if ((rc = launch_step(c, nullptr))) {
          hipGraph_t dummy = nullptr;
          (void)hipStreamEndCapture(c->stream, &dummy);
          if (dummy) (void)hipGraphDestroy(dummy);
          return rc;
        }
      }
      HIPCHK(c, hipStreamEndCapture(c->stream, &g));
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(c, INU_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      it = c->graphs.emplace(m, ex).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second, c->stream));
    left -= m;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_steps_timed(inu_ctx* c, uint32_t n, float* ms_out3) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!ms_out3) return fail(c, INU_EINVAL, "null output");
  if (n == 0 || n > 100000u) return fail(c, INU_EINVAL, "inu_steps_timed: n must be in [1, 100000]");
  const size_t need = 3 * (size_t)n;
  while (c->events.size() < need) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->events.push_back(e);
  }
  for (uint32_t k = 0; k < n; ++k)
    if ((rc = launch_step(c, &c->events[3 * (size_t)k]))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double t0 = 0.0, t1 = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    float a0 = 0.0f, a1 = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&a0, c->events[3 * (size_t)k], c->events[3 * (size_t)k + 1]));
    HIPCHK(c, hipEventElapsedTime(&a1, c->events[3 * (size_t)k + 1], c->events[3 * (size_t)k + 2]));
    t0 += a0;
    t1 += a1;
  }
  float tot = 0.0f;
  HIPCHK(c, hipEventElapsedTime(&tot, c->events[0], c->events[3 * (size_t)(n - 1) + 2]));
  ms_out3[0] = (float)t0;
  ms_out3[1] = (float)t1;
  ms_out3[2] = tot;
  return 0;
}

int inu_download(inu_ctx* c, int which, void* out) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out) return fail(c, INU_EINVAL, "null output");
  const void* src = nullptr;
  size_t bytes = 0;  // 0: a cell field
  switch (which) {
    case INU_FIELD_H: src = c->a.h; break;
    case INU_FIELD_HU: src = c->a.hu; break;
    case INU_FIELD_HV: src = c->a.hv; break;
    case INU_FIELD_H_LO: src = c->a.hlo; break;
    case INU_FIELD_H1: src = c->a.h1; break;
    case INU_FIELD_HU1: src = c->a.hu1; break;
    case INU_FIELD_HV1: src = c->a.hv1; break;
    case INU_FIELD_H1_LO: src = c->a.h1lo; break;
    case INU_FIELD_CLAMP: src = c->a.clamp; break;
    case INU_FIELD_BLOCK_WET: src = c->a.wet; bytes = (size_t)c->NB * 4; break;
    case INU_FIELD_BLOCK_ACTIVE: src = c->a.active; bytes = (size_t)c->NB * 4; break;
    case INU_FIELD_BLOCK_FORCED: src = c->a.forced; bytes = (size_t)c->NB * 4; break;
    default: return fail(c, INU_EINVAL, "unknown field");
  }
  if (bytes == 0) return download_cells(c, static_cast<const float*>(src), static_cast<float*>(out));
  HIPCHK(c, hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_upload(inu_ctx* c, int which, const void* in) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!in) return fail(c, INU_EINVAL, "null input");
  void* dst = nullptr;
  switch (which) {
    case INU_FIELD_H: dst = c->a.h; break;
    case INU_FIELD_HU: dst = c->a.hu; break;
    case INU_FIELD_HV: dst = c->a.hv; break;
    case INU_FIELD_H_LO: dst = c->a.hlo; break;
    case INU_FIELD_H1: dst = c->a.h1; break;
    case INU_FIELD_HU1: dst = c->a.hu1; break;
    case INU_FIELD_HV1: dst = c->a.hv1; break;
    case INU_FIELD_H1_LO: dst = c->a.h1lo; break;
    case INU_FIELD_CLAMP: dst = c->a.clamp; break;
    default: return fail(c, INU_EINVAL, "field cannot be uploaded");
  }
  if ((rc = upload_cells(c, static_cast<float*>(dst), static_cast<const float*>(in)))) return rc;
  return which == INU_FIELD_H ? depths_replaced(c) : 0;
}

int inu_probe(inu_ctx* c, uint32_t n, const float* pairs, float* out7, float g, float dry_tol) {
  if (!c) return INU_EINVAL;
  if (!pairs || !out7) return fail(c, INU_EINVAL, "null array");
  if (n == 0) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  float *din = nullptr, *dout = nullptr;
  HIPCHK(c, hipMalloc(&din, (size_t)n * 8));
  hipError_t e = hipMalloc(&dout, (size_t)n * 4 * kProbeOut);
  if (e != hipSuccess) {
    (void)hipFree(din);
    return fail(c, INU_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  e = hipMemcpyAsync(din, pairs, (size_t)n * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, c->stream, din, dout, n, make_phys(g, dry_tol, 0.8f));
    e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out7, dout, (size_t)n * 4 * kProbeOut, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(din);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("probe: ") + hipGetErrorString(e));
  return 0;
}

int inu_probe_rcp(inu_ctx* c, uint32_t lo, uint32_t hi, uint64_t* mismatches, uint32_t* first) {
  if (!c) return INU_EINVAL;
  if (!mismatches || !first) return fail(c, INU_EINVAL, "inu_probe_rcp: null output");
  if (lo > hi || hi > 0x80000000u) return fail(c, INU_EINVAL, "inu_probe_rcp: need lo <= hi <= 0x80000000");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long* dcnt = nullptr;
  uint32_t* dfirst = nullptr;
  HIPCHK(c, hipMalloc(&dcnt, sizeof(unsigned long long)));
  hipError_t e = hipMalloc(&dfirst, sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipFree(dcnt);
    return fail(c, INU_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  const unsigned long long zero = 0ull;
  const uint32_t none = 0xffffffffu;
  e = hipMemcpyAsync(dcnt, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dfirst, &none, sizeof(none), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && hi > lo) {
    hipLaunchKernelGGL(k_probe_rcp, dim3(8192), dim3(256), 0, c->stream, lo, hi, dcnt, dfirst);
    e = hipGetLastError();
  }
  unsigned long long cnt = 0ull;
  if (e == hipSuccess) e = hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(first, dfirst, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dcnt);
  (void)hipFree(dfirst);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("probe_rcp: ") + hipGetErrorString(e));
  *mismatches = cnt;
  return 0;
}

int inu_probe_clock(inu_ctx* c, int mode, inu_ctrl* ctrl, uint32_t* dt4) {
  if (!c) return INU_EINVAL;
  if (!ctrl || !dt4) return fail(c, INU_EINVAL, "inu_probe_clock: null control block or dt words");
  if (mode < 0 || mode > 2) return fail(c, INU_EINVAL, "inu_probe_clock: mode must be 0 (start), 1 (step) or 2 (retarget)");
  HIPCHK(c, hipSetDevice(c->device));
  // A scenario record and control words of their own: the context's mesh,
  // lists and clock are not touched.
  Scen sc{};
  sc.ctrl = *ctrl;
  sc.dt_cur = dt4[0];
  sc.dt_next = dt4[1];
  Ctl ctl{};
  void* dev = nullptr;
  HIPCHK(c, hipMalloc(&dev, sizeof(Scen) + sizeof(Ctl)));
  Scen* dscen = static_cast<Scen*>(dev);
  Ctl* dctl = reinterpret_cast<Ctl*>(dscen + 1);
  hipError_t e = hipMemcpyAsync(dscen, &sc, sizeof(Scen), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dctl, &ctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    Args a{};
    a.scen = dscen;
    a.ctl = dctl;
    a.n_scen = 1;
    a.n_dwg = 0;
    hipLaunchKernelGGL(k_step_control, dim3(1), dim3(256), 0, c->stream, a, mode);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&sc, dscen, sizeof(Scen), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("inu_probe_clock: ") + hipGetErrorString(e));
  *ctrl = sc.ctrl;
  dt4[0] = sc.dt_cur;
  dt4[1] = sc.dt_next;
  dt4[2] = 0u;
  dt4[3] = 0u;
  return 0;
}

int inu_probe_open(inu_ctx* c, uint32_t n, const float* cases6, const uint32_t* kinds, const float* values,
                   float* out4, float g, float dry_tol) {
  if (!c) return INU_EINVAL;
  if (!cases6 || !kinds || !values || !out4) return fail(c, INU_EINVAL, "null array");
  if (n == 0) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t in_b = (size_t)n * 24, k_b = (size_t)n * 4, out_b = (size_t)n * 16;
  char* buf = nullptr;
  HIPCHK(c, hipMalloc(&buf, in_b + 2 * k_b + out_b));
  float* din = reinterpret_cast<float*>(buf);
  uint32_t* dk = reinterpret_cast<uint32_t*>(buf + in_b);
  float* dv = reinterpret_cast<float*>(buf + in_b + k_b);
  float* dout = reinterpret_cast<float*>(buf + in_b + 2 * k_b);
  hipError_t e = hipMemcpyAsync(din, cases6, in_b, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dk, kinds, k_b, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, values, k_b, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_probe_open, dim3((n + 255) / 256), dim3(256), 0, c->stream, din, dk, dv, dout, n,
                       make_phys(g, dry_tol, 0.8f));
    e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, out_b, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("probe_open: ") + hipGetErrorString(e));
  return 0;
}

int inu_stream_copy(inu_ctx* c, uint64_t n_bytes, uint32_t reps, float* ms_out) {
  if (!c) return INU_EINVAL;
  if (!ms_out || reps == 0u || n_bytes == 0u || n_bytes % 16u != 0u)
    return fail(c, INU_EINVAL, "inu_stream_copy: n_bytes must be a positive multiple of 16, reps positive");
  HIPCHK(c, hipSetDevice(c->device));
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipMalloc(&buf[0], n_bytes);
  if (e == hipSuccess) e = hipMalloc(&buf[1], n_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(buf[0], 0, n_bytes, c->stream);
  if (e == hipSuccess) e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  const size_t n = n_bytes / 16u;
  const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)c->cus * 64u);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (const float4*)buf[0], (float4*)buf[1], n);
    e = hipEventRecord(ev[0], c->stream);
  }
  for (uint32_t r = 0; e == hipSuccess && r < reps; ++r)
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (const float4*)buf[0], (float4*)buf[1], n);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  for (void* q : buf)
    if (q) (void)hipFree(q);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("stream copy: ") + hipGetErrorString(e));
  *ms_out = ms / (float)reps;
  return 0;
}

int inu_stream_bw(inu_ctx* c, uint64_t n_bytes, uint32_t reps, int kind, uint32_t unroll, uint32_t wg_per_cu,
                  float* ms_out) {
  if (!c) return INU_EINVAL;
  if (!ms_out || reps == 0u || n_bytes == 0u || n_bytes % 16u != 0u)
    return fail(c, INU_EINVAL, "inu_stream_bw: n_bytes must be a positive multiple of 16, reps positive");
  if (kind < 0 || kind > 3) return fail(c, INU_EINVAL, "inu_stream_bw: kind must be 0 (copy), 1 (read), 2 (write), 3 (hipMemcpy)");
  if (unroll != 1u && unroll != 2u && unroll != 4u && unroll != 8u)
    return fail(c, INU_EINVAL, "inu_stream_bw: unroll must be 1, 2, 4 or 8");
  if (wg_per_cu == 0u || wg_per_cu > 32u) return fail(c, INU_EINVAL, "inu_stream_bw: wg_per_cu in [1, 32]");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = n_bytes / 16u;
  const uint32_t grid = (uint32_t)std::max<size_t>(1, std::min<size_t>((size_t)c->cus * wg_per_cu, (n + 255) / 256));
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipMalloc(&buf[0], n_bytes);
  if (e == hipSuccess) e = hipMalloc(&buf[1], std::max<size_t>(n_bytes, (size_t)grid * 256u * 4u));
  if (e == hipSuccess) e = hipMemsetAsync(buf[0], 0, n_bytes, c->stream);
  if (e == hipSuccess) e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  const f32x4* in = static_cast<const f32x4*>(buf[0]);
  f32x4* out = static_cast<f32x4*>(buf[1]);
  auto launch = [&]() -> hipError_t {
    if (kind == 3) return hipMemcpyAsync(buf[1], buf[0], n_bytes, hipMemcpyDeviceToDevice, c->stream);
#define INU_BW_LAUNCH(UU)                                                                                  \
  if (kind == 0) hipLaunchKernelGGL(k_bw_copy<UU>, dim3(grid), dim3(256), 0, c->stream, in, out, n);      \
  else if (kind == 1) hipLaunchKernelGGL(k_bw_read<UU>, dim3(grid), dim3(256), 0, c->stream, in, (float*)out, n); \
  else hipLaunchKernelGGL(k_bw_write<UU>, dim3(grid), dim3(256), 0, c->stream, out, n);
    if (unroll == 1u) { INU_BW_LAUNCH(1) }
    else if (unroll == 2u) { INU_BW_LAUNCH(2) }
    else if (unroll == 4u) { INU_BW_LAUNCH(4) }
    else { INU_BW_LAUNCH(8) }
#undef INU_BW_LAUNCH
    return hipGetLastError();
  };
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(ev[0], c->stream);
  for (uint32_t r = 0; e == hipSuccess && r < reps; ++r) e = launch();
  if (e == hipSuccess) e = hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  for (void* q : buf)
    if (q) (void)hipFree(q);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("inu_stream_bw: ") + hipGetErrorString(e));
  *ms_out = ms / (float)reps;
  return 0;
}

int inu_autotune(inu_ctx* c, uint32_t reps, uint32_t* out4) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out4 || reps == 0u || reps > 1000u) return fail(c, INU_EINVAL, "inu_autotune: reps in [1, 1000] and an output");
  uint32_t any_dt = 0;
  if ((rc = copy_sync(c, &any_dt, &c->ctl->any_dt, 4, hipMemcpyDeviceToHost))) return rc;
  if (any_dt == 0u) return fail(c, INU_ESTATE, "inu_autotune: no scenario has a step to take (call after inu_prologue)");
  const bool comp = (c->flags & INU_FLAG_COMPENSATED) != 0;
  const bool fm = (c->flags & INU_FLAG_EXACT_ORDER) == 0;
  // The predictor writes only the predictor state, which the next step
  // rewrites before reading; the clamp tally stays off while timing.
  Args a = c->a;
  a.track_clamp = 0u;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  float best_ms = 0.0f;
  int best = 0;
  for (int tl = 0; tl < kTiles && e == hipSuccess; ++tl) {
    launch_stage<0>(c, a, comp, fm, tl);
    e = hipEventRecord(ev[0], c->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) launch_stage<0>(c, a, comp, fm, tl);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ev[1], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    out4[1 + tl] = (uint32_t)std::min(4.0e9f, ms * 1000.0f / (float)reps);  // microseconds per launch
    if (tl == 0 || ms < best_ms) {
      best_ms = ms;
      best = tl;
    }
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("autotune: ") + hipGetErrorString(e));
  c->tile = best;
  out4[0] = (uint32_t)best;
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  return 0;
}

int inu_set_tile(inu_ctx* c, uint32_t tile) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (tile >= (uint32_t)kTiles) return fail(c, INU_EINVAL, "tile level out of range");
  c->tile = (int)tile;
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  return 0;
}

static bool rc_nonneg_required(const inu_ctx* c) { return !c->src_nonneg; }

int inu_set_source_rates(inu_ctx* c, uint32_t n, const uint32_t* cell, const float* rate) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n == 0) return 0;
  if (!cell || !rate) return fail(c, INU_EINVAL, "inu_set_source_rates: null array");
  if (n > c->n_src) return fail(c, INU_EINVAL, "more rates than source cells");
  std::vector<uint32_t> ic(n);
  for (uint32_t k = 0; k < n; ++k) {
    if (cell[k] >= c->N) return fail(c, INU_EINVAL, "source cell out of range");
    if (!std::isfinite(rate[k])) return fail(c, INU_EINVAL, "source rate must be finite");
    ic[k] = internal_cell(c, cell[k]);
    if (!std::binary_search(c->src_cells.begin(), c->src_cells.end(), ic[k]))
      return fail(c, INU_EINVAL, "inu_set_source_rates: a cell is not a source cell (use inu_set_sources)");
  }
  uint32_t* dcell = nullptr;
  float* drate = nullptr;
  hipError_t e = hipMallocAsync((void**)&dcell, (size_t)n * 4, c->stream);
  if (e == hipSuccess) e = hipMallocAsync((void**)&drate, (size_t)n * 4, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dcell, ic.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(drate, rate, (size_t)n * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_scatter_rates, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->a.src, dcell, drate, n);
    e = hipGetLastError();
  }
  if (dcell) (void)hipFreeAsync(dcell, c->stream);
  if (drate) (void)hipFreeAsync(drate, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("set_source_rates: ") + hipGetErrorString(e));
  for (uint32_t k = 0; k < n; ++k)
    if (rate[k] < 0.0f) c->src_nonneg = false;
  if (!c->src_nonneg && c->a.src_neg == 0u) {
    // The stage kernels' all-wet test changes: re-record the step graphs.
    c->a.src_neg = 1u;
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
  }
  if (c->adj.on) {
    // The recording replays these rates from the current step on.
    if (rc_nonneg_required(c)) return fail(c, INU_EINVAL, "the adjoint needs non-negative source rates");
    std::vector<Scen> sc;
    if ((rc = scen_read(c, sc))) return rc;
    const uint32_t step = (uint32_t)sc[0].ctrl.steps;
    const size_t NP = (size_t)c->NB * kBlockCells;
    float* ev = nullptr;
    if (!c->adj.events.empty() && c->adj.events.back().first == step) {
      ev = c->adj.events.back().second;
    } else {
      if ((rc = dalloc(c, &ev, NP))) return rc;
      c->adj.events.emplace_back(step, ev);
    }
    if ((rc = copy_sync(c, ev, c->a.src, NP * 4, hipMemcpyDeviceToDevice))) return rc;
  }
  return 0;
}

int inu_set_source_rates_noalloc(inu_ctx* c, uint32_t n, const uint32_t* cell, const float* rate) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n == 0) return 0;
  if (!cell || !rate) return fail(c, INU_EINVAL, "inu_set_source_rates_noalloc: null array");
  if (n > kScratchRates)
    return fail(c, INU_EINVAL, "inu_set_source_rates_noalloc: at most 65536 rates per call (split the list)");
  if (n > c->n_src) return fail(c, INU_EINVAL, "more rates than source cells");
  if (c->adj.on)
    return fail(c, INU_ESTATE,
                "inu_set_source_rates_noalloc: not during an adjoint recording, whose rate log allocates "
                "(use inu_set_source_rates)");
  void* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!base) {
    const hipError_t e = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_err[c->device] : hipErrorInvalidDevice;
    return fail(c, INU_EHIP, std::string("inu_set_source_rates_noalloc: no scratch address on this device: ") +
                                 hipGetErrorString(e));
  }
  for (uint32_t k = 0; k < n; ++k) {
    if (cell[k] >= c->N) return fail(c, INU_EINVAL, "source cell out of range");
    if (!std::isfinite(rate[k])) return fail(c, INU_EINVAL, "source rate must be finite");
    // A negative rate among non-negative sources would change the stage
    // kernels' all-wet test and need the step graphs recorded again.
    if (rate[k] < 0.0f && c->src_nonneg)
      return fail(c, INU_EINVAL,
                  "inu_set_source_rates_noalloc: a negative rate among non-negative sources re-records the step "
                  "graphs (use inu_set_source_rates)");
    if (!std::binary_search(c->src_cells.begin(), c->src_cells.end(), internal_cell(c, cell[k])))
      return fail(c, INU_EINVAL, "inu_set_source_rates_noalloc: a cell is not a source cell (use inu_set_sources)");
  }
  HIPCHK(c, hipSetDevice(c->device));
  uint32_t* dcell = static_cast<uint32_t*>(base);
  float* drate = reinterpret_cast<float*>(dcell + kScratchRates);
  // Internal cell ids through a stack buffer, one chunk at a time (each copy
  // is waited for before the buffer is refilled).
  uint32_t chunk[1024];
  hipError_t e = hipSuccess;
  for (uint32_t k0 = 0; e == hipSuccess && k0 < n; k0 += 1024u) {
    const uint32_t m = std::min(1024u, n - k0);
    for (uint32_t k = 0; k < m; ++k) chunk[k] = internal_cell(c, cell[k0 + k]);
    e = hipMemcpyAsync(dcell + k0, chunk, (size_t)m * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(drate, rate, (size_t)n * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_scatter_rates, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->a.src, dcell, drate, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("set_source_rates_noalloc: ") + hipGetErrorString(e));
  return 0;
}

int inu_set_open_stretches(inu_ctx* c, uint32_t n_stretches, const uint32_t* stretch_ptr) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->n_open == 0) return fail(c, INU_EINVAL, "the mesh has no open edges");
  if (n_stretches == 0 || !stretch_ptr) return fail(c, INU_EINVAL, "inu_set_open_stretches: no stretches");
  if (n_stretches > c->n_open) return fail(c, INU_EINVAL, "more stretches than open edges");
  if (stretch_ptr[0] != 0 || stretch_ptr[n_stretches] != c->n_open)
    return fail(c, INU_EINVAL, "stretch_ptr must run from 0 to the open edge count");
  for (uint32_t s = 0; s < n_stretches; ++s)
    if (stretch_ptr[s + 1] <= stretch_ptr[s]) return fail(c, INU_EINVAL, "every stretch needs an open edge");
  if ((rc = upload(c, c->stretch_ptr, stretch_ptr, n_stretches + 1))) return rc;
  c->n_stretch = n_stretches;
  HIPCHK(c, hipMemsetAsync(c->tally, 0, 4 * (size_t)n_stretches * 4, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  return 0;
}

int inu_set_open_conditions(inu_ctx* c, uint32_t n, const uint32_t* kinds, const float* values) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n != c->n_open) return fail(c, INU_EINVAL, "one condition per open edge");
  if (n == 0) return 0;
  if (!kinds || !values) return fail(c, INU_EINVAL, "inu_set_open_conditions: null array");
  for (uint32_t k = 0; k < n; ++k) {
    if (kinds[k] > INU_OPEN_INFLOW) return fail(c, INU_EINVAL, "unknown open boundary condition");
    if (!std::isfinite(values[k])) return fail(c, INU_EINVAL, "open boundary values must be finite");
  }
  if ((rc = upload(c, const_cast<uint32_t*>(c->open_h.kind), kinds, n))) return rc;
  if ((rc = upload(c, const_cast<float*>(c->open_h.value), values, n))) return rc;
  // A block with an edge that can bring water in from outside counts as wet
  // while these conditions hold (reference gpu_resident_batch.py:447-458).
  std::vector<uint32_t> forced(c->NB, 0u);
  for (uint32_t k = 0; k < n; ++k)
    if ((kinds[k] == INU_OPEN_INFLOW || kinds[k] == INU_OPEN_STAGE) && values[k] > 0.0f)
      forced[c->open_cell_h[k] / kBlockCells] = 1u;
  return upload(c, c->a.forced, forced.data(), forced.size());
}

int inu_open_volumes(inu_ctx* c, uint32_t n_stretches, double* out2) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n_stretches != c->n_stretch)
    return fail(c, INU_EINVAL, "inu_open_volumes: n_stretches is not the stretch count");
  if (n_stretches == 0) return 0;
  if (!out2) return fail(c, INU_EINVAL, "null output");
  std::vector<float> t(4 * (size_t)c->n_stretch);
  if ((rc = copy_sync(c, t.data(), c->tally, t.size() * 4, hipMemcpyDeviceToHost))) return rc;
  for (uint32_t s = 0; s < c->n_stretch; ++s) {
    out2[2 * s] = (double)t[4 * s] + (double)t[4 * s + 1];
    out2[2 * s + 1] = (double)t[4 * s + 2] + (double)t[4 * s + 3];
  }
  return 0;
}

int inu_set_bed(inu_ctx* c, const float* z) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!z) return fail(c, INU_EINVAL, "inu_set_bed: null array");
  std::vector<float> zi((size_t)c->NB * kBlockCells, 0.0f);
  for (uint32_t s = 0; s < c->n_scen; ++s)
    std::copy(z + c->cell0[s], z + c->cell0[s + 1], zi.begin() + (size_t)c->block0[s] * kBlockCells);
  for (uint32_t i = 0; i < c->N; ++i)
    if (!std::isfinite(z[i])) return fail(c, INU_EINVAL, "bed elevation must be finite");
  // Face on the higher bed, steps in FP64 then FP32 (reference mesh/geometry.py:241-254).
  for (size_t q = 0; q < c->stream_edge.size(); ++q) {
    const uint32_t e = c->stream_edge[q];
    const int32_t r = c->e_cr[e];
    if (r < 0) {  // a wall or an open edge: no bed step
      c->h_ldz[q] = 0.0f;
      continue;
    }
    const double zl = zi[c->e_cl[e]], zr = zi[r];
    const double zf = std::max(zl, zr);
    const float dzl = (float)(zf - zl), dzr = (float)(zf - zr);
    c->h_ldz[q] = (dzl > 0.0f) ? dzl : -dzr;
  }
  return upload(c, const_cast<float*>(c->a.ldz), c->h_ldz.data(), c->h_ldz.size());
}

}  // extern "C"

namespace {

// Frees a recording's buffers (log, lists, gauges, series, checkpoints,
// source events, reverse-pass work) and every cached graph.
static void adj_release(inu_ctx* c) {
  auto& d = c->adj;
  auto drop = [&](void* q) {
    auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
    if (q && it != c->allocs.end()) {
      (void)hipFree(q);
      c->allocs.erase(it);
    }
  };
  for (void* q : {(void*)d.log, (void*)d.codes, (void*)d.gauges, (void*)d.series, (void*)d.ck, (void*)d.cur,
                  (void*)d.tape, (void*)d.lam, (void*)d.l1, (void*)d.gn, (void*)d.dseed, (void*)d.rseries,
                  (void*)d.gdz})
    drop(q);
  for (auto& ev : d.events) drop((void*)ev.second);
  d.events.clear();
  d.log = nullptr;
  d.codes = nullptr;
  d.gauges = nullptr;
  d.series = nullptr;
  d.ck = nullptr;
  d.cur = nullptr;
  d.tape = d.lam = d.l1 = d.gn = d.dseed = d.rseries = nullptr;
  d.gdz = nullptr;
  d.on = false;
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

// Graph keys of the adjoint's step blocks (length in the low 28 bits).
constexpr uint32_t kKeyRecord = 0x80000000u, kKeyReplayRec = 0x40000000u, kKeyReplay = 0x20000000u,
                   kKeyReverse = 0x10000000u, kKeyLenMask = 0x0fffffffu;

// Runs enqueue(m) from a graph cached under key | m, captured on first use.
template <class F>
static int run_graph(inu_ctx* c, uint32_t key, uint32_t m, F enqueue) {
  auto it = c->graphs.find(key | m);
  if (it == c->graphs.end()) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(m);
    const hipError_t e_end = hipStreamEndCapture(c->stream, &g);
    if (rc || e_end != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(c, INU_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e_end));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, INU_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    it = c->graphs.emplace(key | m, ex).first;
  }
  HIPCHK(c, hipGraphLaunch(it->second, c->stream));
  return 0;
}

// Default thin-layer depth of the reverse pass, in dry_tol: 1 cm at the usual
// 1e-4 m (on the 250 k-cell inflow flood the gradient's distance to central
// differences is smallest near there).
constexpr float kAdjThinDry = 100.0f;

static float adj_thin_depth(const inu_ctx* c) {
  return c->adj_thin < 0.0f ? kAdjThinDry * c->p.dry : c->adj_thin;
}

static AdjStep adj_step_args(inu_ctx* c, bool reverse) {
  const auto& d = c->adj;
  return AdjStep{d.codes, d.log, d.cur, reverse ? d.dseed : nullptr, d.gauges, d.lam, d.n_gauges, d.every,
                 d.max_samples};
}

// m recorded steps replayed from the cursor on: each step's lists and dt from
// the recording, the two stages (writing the tape when rec), the idle
// blocks' sources; series != null samples the gauges.
static int adj_enqueue_replay(inu_ctx* c, uint32_t m, bool rec, float* series) {
  const auto& d = c->adj;
  Args ar = c->a;
  ar.tape = d.tape;
  ar.adj_cur = d.cur;
  const AdjStep st = adj_step_args(c, false);
  const bool comp = (c->flags & INU_FLAG_COMPENSATED) != 0;
  const bool fm = (c->flags & INU_FLAG_EXACT_ORDER) == 0;
  const uint32_t gl = (c->NB + 255) / 256;
  for (uint32_t j = 0; j < m; ++j) {
    hipLaunchKernelGGL(k_adj_lists, dim3(gl), dim3(256), 0, c->stream, ar, st);
    if (rec) {
      launch_stage_rec<0>(c, ar, comp, fm);
      launch_stage_rec<1>(c, ar, comp, fm);
    } else {
      launch_stage<0>(c, ar, comp, fm, c->tile);
      launch_stage<1>(c, ar, comp, fm, c->tile);
    }
    if (c->n_src) launch_source_idle(c, ar, comp, fm);
    hipLaunchKernelGGL(k_adj_end, dim3(1), dim3(256), 0, c->stream, ar, d.cur, 1, series, d.gauges, d.n_gauges,
                       d.every, d.max_samples);
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

// m recorded steps reversed from the cursor down.
static int adj_enqueue_reverse(inu_ctx* c, uint32_t m) {
  const auto& d = c->adj;
  AdjArgs x{c->a, d.tape, d.log, d.cur, d.lam, d.l1, d.gn, d.gdz};
  x.a.p.adj_thin = adj_thin_depth(c);
  const AdjStep st = adj_step_args(c, true);
  const uint32_t gl = (c->NB + 255) / 256;
  for (uint32_t j = 0; j < m; ++j) {
    hipLaunchKernelGGL(k_adj_lists, dim3(gl), dim3(256), 0, c->stream, c->a, st);
    hipLaunchKernelGGL(k_adj_stage<1>, dim3(d.grid_adj), dim3(256), c->lds_full, c->stream, x);
    hipLaunchKernelGGL(k_adj_stage<0>, dim3(d.grid_adj), dim3(256), c->lds_full, c->stream, x);
    hipLaunchKernelGGL(k_adj_end, dim3(1), dim3(256), 0, c->stream, c->a, d.cur, -1, nullptr, d.gauges, d.n_gauges,
                       d.every, d.max_samples);
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

// The source rates in effect at step k: the last event at or before it.
static const float* adj_rates_at(const inu_ctx* c, uint32_t k) {
  const float* r = nullptr;
  for (const auto& ev : c->adj.events)
    if (ev.first <= k) r = ev.second;
  return r;
}

// Replays steps [ka, kb) of the segment starting at k0 from the state in
// place (the state of step ka, with the rates in effect at ka), changing the
// rates at each recorded event.
static int adj_replay_range(inu_ctx* c, uint32_t ka, uint32_t kb, uint32_t k0, bool rec, float* series) {
  auto& d = c->adj;
  const size_t NP = (size_t)c->NB * kBlockCells;
  hipLaunchKernelGGL(k_adj_cursor, dim3(1), dim3(1), 0, c->stream, c->a, d.cur, ka, k0);
  uint32_t k = ka;
  while (k < kb) {
    uint32_t e = kb;
    for (const auto& ev : d.events)
      if (ev.first > k && ev.first < e) e = ev.first;
    int rc = run_graph(c, rec ? kKeyReplayRec : kKeyReplay, e - k,
                       [&](uint32_t m) { return adj_enqueue_replay(c, m, rec, series); });
    if (rc) return rc;
    k = e;
    if (k < kb) HIPCHK(c, hipMemcpyAsync(c->a.src, adj_rates_at(c, k), NP * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  return 0;
}

// Restores the state of step k from the checkpoint at ck, and the rates in
// effect then.
static int adj_restore_from(inu_ctx* c, const float* ck, uint32_t k) {
  const size_t NP = (size_t)c->NB * kBlockCells;
  const Args& a = c->a;
  HIPCHK(c, hipMemcpyAsync(a.h, ck, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(a.part_dry, 1, (size_t)c->NB * 4, c->stream));
  HIPCHK(c, hipMemcpyAsync(a.hu, ck + NP, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(a.hv, ck + 2 * NP, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(a.hlo, ck + 3 * NP, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  const float* rates = adj_rates_at(c, k);
  if (rates) HIPCHK(c, hipMemcpyAsync(a.src, rates, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// Restores checkpoint cs of the uniform scheme (the state of step cs * seg).
static int adj_restore(inu_ctx* c, uint32_t cs) {
  const size_t NP = (size_t)c->NB * kBlockCells;
  return adj_restore_from(c, c->adj.ck + (size_t)cs * 4 * NP, cs * c->adj.seg);
}

// Writes the state in place to the checkpoint at ck.
static int adj_store_to(inu_ctx* c, float* ck) {
  const size_t NP = (size_t)c->NB * kBlockCells;
  const Args& a = c->a;
  HIPCHK(c, hipMemcpyAsync(ck, a.h, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ck + NP, a.hu, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ck + 2 * NP, a.hv, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ck + 3 * NP, a.hlo, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// Largest segment count the binomial schedule takes (its table is built in
// O(n^2 s) host operations).
constexpr uint32_t kMaxScheduleSegments = 2048u, kMaxSlots = 64u;

// The optimal offline checkpoint schedule for reversing n segments with the
// start state held and `free` more checkpoint slots (Griewank's binomial
// checkpointing, built by dynamic programming over the recursion it
// solves). cost[l * (free + 1) + f]: the fewest segments advanced without
// tape to reverse l segments with f free slots; split[...]: how many
// segments to advance before storing the next checkpoint. Reversing one
// segment costs no advance (its replay with tape is counted apart, once per
// segment); with no free slot, segment i of l is reached by advancing i.
struct Schedule {
  uint32_t free = 0;
  std::vector<unsigned long long> cost;
  std::vector<uint32_t> split;
  unsigned long long at(uint32_t l, uint32_t f) const { return cost[(size_t)l * (free + 1) + f]; }
  uint32_t cut(uint32_t l, uint32_t f) const { return split[(size_t)l * (free + 1) + f]; }
};

static Schedule revolve_schedule(uint32_t n, uint32_t free) {
  Schedule t;
  t.free = free;
  const size_t w = (size_t)free + 1;
  t.cost.assign((size_t)(n + 1) * w, 0ull);
  t.split.assign((size_t)(n + 1) * w, 0u);
  for (uint32_t l = 2; l <= n; ++l) {
    t.cost[l * w] = (unsigned long long)l * (l - 1) / 2;
    for (uint32_t f = 1; f <= free; ++f) {
      unsigned long long best = ~0ull;
      uint32_t arg = 1;
      for (uint32_t j = 1; j < l; ++j) {
        const unsigned long long v = j + t.cost[(l - j) * w + f - 1] + t.cost[j * w + f];
        if (v < best) {
          best = v;
          arg = j;
        }
      }
      t.cost[l * w + f] = best;
      t.split[l * w + f] = arg;
    }
  }
  return t;
}

// Replays segment i (the state of its first step in place) writing the tape,
// then reverses it.
static int adj_turn(inu_ctx* c, uint32_t i, uint32_t K) {
  auto& d = c->adj;
  const uint32_t k0 = i * d.seg, k1 = std::min(K, k0 + d.seg);
  int rc = adj_replay_range(c, k0, k1, k0, true, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_adj_cursor, dim3(1), dim3(1), 0, c->stream, c->a, d.cur, k1 - 1, k0);
  rc = run_graph(c, kKeyReverse, k1 - k0, [&](uint32_t m) { return adj_enqueue_reverse(c, m); });
  d.replayed[1] += k1 - k0;
  return rc;
}

// The backward pass by the binomial schedule: K recorded steps in segments of
// seg steps, checkpoint slot 0 holding the start.
static int adj_backward_binomial(inu_ctx* c, uint32_t K) {
  auto& d = c->adj;
  const size_t NP = (size_t)c->NB * kBlockCells;
  const uint32_t n = (K + d.seg - 1) / d.seg;
  const Schedule sch = revolve_schedule(n, std::min(d.n_ck - 1, n - 1));
  std::vector<uint8_t> busy(d.n_ck, 0);
  busy[0] = 1;
  auto slot = [&](uint32_t t) { return d.ck + (size_t)t * 4 * NP; };
  // The state of segment q, from the state of segment a in slot t, one
  // segment per replay (so the replay graphs are those of one segment).
  auto advance = [&](uint32_t t, uint32_t a, uint32_t q) -> int {
    int rc = adj_restore_from(c, slot(t), a * d.seg);
    for (uint32_t p = a; p < q && !rc; ++p) {
      rc = adj_replay_range(c, p * d.seg, (p + 1) * d.seg, p * d.seg, false, nullptr);
      d.replayed[0] += d.seg;
      // A range applies the events strictly inside it: leave the rates in
      // effect at the next segment's first step (an event there included),
      // as a restore at that step would.
      const float* rates = adj_rates_at(c, (p + 1) * d.seg);
      if (!rc && rates) HIPCHK(c, hipMemcpyAsync(c->a.src, rates, NP * 4, hipMemcpyDeviceToDevice, c->stream));
    }
    return rc;
  };
  // Reverses segments [a, b), the state of segment a in slot t, f slots free.
  std::function<int(uint32_t, uint32_t, uint32_t, uint32_t)> rev = [&](uint32_t a, uint32_t b, uint32_t t,
                                                                      uint32_t f) -> int {
    int rc = 0;
    if (b - a == 1 || f == 0) {
      for (uint32_t i = b; i-- > a && !rc;)
        if (!(rc = advance(t, a, i))) rc = adj_turn(c, i, K);
      return rc;
    }
    const uint32_t m = a + sch.cut(b - a, f);
    uint32_t u = 1;
    while (busy[u]) ++u;
    if ((rc = advance(t, a, m)) || (rc = adj_store_to(c, slot(u)))) return rc;
    busy[u] = 1;
    rc = rev(m, b, u, f - 1);
    busy[u] = 0;
    return rc ? rc : rev(a, m, t, f);
  };
  return rev(0, n, 0, sch.free);
}

}  // namespace

extern "C" {

int inu_adj_begin(inu_ctx* c, uint32_t max_steps, uint32_t seg, uint32_t n_gauges, const uint32_t* gauges,
                  uint32_t every) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->n_scen != 1) return fail(c, INU_EINVAL, "the adjoint records one scenario");
  if (c->n_open) return fail(c, INU_EINVAL, "the adjoint does not differentiate open boundaries");
  if (!(c->flags & INU_FLAG_DETERMINISTIC))
    return fail(c, INU_EINVAL, "the adjoint replays the fixed-order gather: set INU_FLAG_DETERMINISTIC");
  if (max_steps == 0 || seg == 0 || every == 0 || n_gauges == 0 || !gauges)
    return fail(c, INU_EINVAL, "inu_adj_begin: steps, segment, sampling interval and gauges must be given");
  if (max_steps > kKeyLenMask) return fail(c, INU_EINVAL, "inu_adj_begin: too many steps");
  if (!c->src_nonneg) return fail(c, INU_EINVAL, "the adjoint needs non-negative source rates");
  std::vector<uint32_t> sorted(gauges, gauges + n_gauges);
  std::sort(sorted.begin(), sorted.end());
  if (sorted.back() >= c->N) return fail(c, INU_EINVAL, "gauge cell out of range");
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return fail(c, INU_EINVAL, "gauge cells must be distinct");
  std::vector<Scen> sc;
  if ((rc = scen_read(c, sc))) return rc;
  if (sc[0].ctrl.steps != 0ull) return fail(c, INU_ESTATE, "start the recording at step 0 (after inu_prologue)");
  const uint32_t n_segments = (max_steps + std::min(seg, max_steps) - 1) / std::min(seg, max_steps);
  if (c->adj_slots && n_segments > kMaxScheduleSegments)
    return fail(c, INU_EINVAL, "inu_adj_begin: the binomial schedule takes at most 2048 segments (a longer seg)");
  adj_release(c);
  auto& d = c->adj;
  d.max_steps = max_steps;
  d.seg = std::min(seg, max_steps);
  d.every = every;
  d.n_gauges = n_gauges;
  d.max_samples = max_steps / every;
  d.slots = c->adj_slots;
  d.n_ck = d.slots ? std::min(d.slots, n_segments) : n_segments + 1;
  const size_t NP = (size_t)c->NB * kBlockCells;
  std::vector<uint32_t> gi(n_gauges);
  for (uint32_t k = 0; k < n_gauges; ++k) gi[k] = internal_cell(c, gauges[k]);
  rc |= dalloc(c, &d.log, max_steps);
  rc |= dalloc(c, &d.codes, (size_t)max_steps * c->NB);
  rc |= dalloc(c, &d.gauges, n_gauges);
  rc |= dalloc(c, &d.series, (size_t)std::max(d.max_samples, 1u) * n_gauges);
  rc |= dalloc(c, &d.ck, (size_t)d.n_ck * 4 * NP);
  rc |= dalloc(c, &d.cur, 1);
  float* ev = nullptr;
  rc |= dalloc(c, &ev, NP);
  if (rc) {
    if (ev) c->adj.events.emplace_back(0u, ev);
    adj_release(c);
    return rc;
  }
  d.events.emplace_back(0u, ev);
  if ((rc = upload(c, d.gauges, gi.data(), gi.size()))) return rc;
  HIPCHK(c, hipMemsetAsync(d.codes, 0, (size_t)max_steps * c->NB, c->stream));
  HIPCHK(c, hipMemsetAsync(d.series, 0, (size_t)std::max(d.max_samples, 1u) * n_gauges * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(d.cur, 0, sizeof(AdjCur), c->stream));
  HIPCHK(c, hipMemcpyAsync(ev, c->a.src, NP * 4, hipMemcpyDeviceToDevice, c->stream));
  const AdjRec r{d.log,         d.codes, d.series,      d.gauges, d.ck, d.max_steps, d.n_gauges, d.every,
                 d.max_samples, d.seg,   d.slots ? 1u : d.n_ck};
  hipLaunchKernelGGL(k_adj_record, dim3(std::min<uint32_t>(c->NB, 1024u)), dim3(256), 0, c->stream, c->a, r);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d.on = true;
  return 0;
}

int inu_adj_steps(inu_ctx* c, uint32_t n) {
  int rc = check_ready(c);
  if (rc) return rc;
  auto& d = c->adj;
  if (!d.on) return fail(c, INU_ESTATE, "no adjoint recording (inu_adj_begin)");
  if (n == 0) return 0;
  if (n > kKeyLenMask) return fail(c, INU_EINVAL, "inu_adj_steps: too many steps at once");
  const AdjRec r{d.log,         d.codes, d.series,      d.gauges, d.ck, d.max_steps, d.n_gauges, d.every,
                 d.max_samples, d.seg,   d.slots ? 1u : d.n_ck};
  rc = run_graph(c, kKeyRecord, n, [&](uint32_t m) {
    for (uint32_t k = 0; k < m; ++k) {
      int r2 = launch_step(c, nullptr);
      if (r2) return r2;
      hipLaunchKernelGGL(k_adj_record, dim3(std::min<uint32_t>(c->NB, 1024u)), dim3(256), 0, c->stream, c->a, r);
    }
    HIPCHK(c, hipGetLastError());
    return 0;
  });
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int inu_adj_series(inu_ctx* c, float* out) {
  int rc = check_ready(c);
  if (rc) return rc;
  auto& d = c->adj;
  if (!d.on) return fail(c, INU_ESTATE, "no adjoint recording (inu_adj_begin)");
  if (!out) return fail(c, INU_EINVAL, "null output");
  return copy_sync(c, out, d.series, (size_t)d.max_samples * d.n_gauges * 4, hipMemcpyDeviceToHost);
  return 0;
}

// Steps recorded so far, each with a nonzero dt.
static int adj_recorded_steps(inu_ctx* c, uint32_t* K) {
  auto& d = c->adj;
  std::vector<Scen> sc;
  int rc = scen_read(c, sc);
  if (rc) return rc;
  if (sc[0].ctrl.steps == 0 || sc[0].ctrl.steps > d.max_steps)
    return fail(c, INU_ESTATE, "recorded steps must be in [1, max_steps]");
  *K = (uint32_t)sc[0].ctrl.steps;
  std::vector<uint32_t> log(*K);
  if ((rc = copy_sync(c, log.data(), d.log, (size_t)*K * 4, hipMemcpyDeviceToHost))) return rc;
  for (uint32_t k = 0; k < *K; ++k)
    if (log[k] == 0u) return fail(c, INU_ESTATE, "a recorded step has dt = 0 (the recording ran past its stop)");
  return 0;
}

// Reverse-pass work buffers, allocated once per recording.
static int adj_work(inu_ctx* c) {
  auto& d = c->adj;
  if (d.tape) return 0;
  const size_t NP = (size_t)c->NB * kBlockCells;
  const size_t L = std::max<size_t>(c->h_ldz.size(), 1);
  int rc = 0;
  rc |= dalloc(c, &d.tape, (size_t)d.seg * kTapeWords * NP);
  rc |= dalloc(c, &d.lam, 3 * NP);
  rc |= dalloc(c, &d.l1, 3 * NP);
  rc |= dalloc(c, &d.gn, NP);
  rc |= dalloc(c, &d.gdz, L);
  rc |= dalloc(c, &d.dseed, (size_t)std::max(d.max_samples, 1u) * d.n_gauges);
  rc |= dalloc(c, &d.rseries, (size_t)std::max(d.max_samples, 1u) * d.n_gauges);
  if (rc) return fail(c, INU_EHIP, "adjoint: out of device memory (the tape needs segment x 56 B per cell)");
  int occ = 0;
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_adj_stage<0>, 256, c->lds_full));
  // A whole multiple of the XCD count (the walk's slices), at least one per XCD.
  d.grid_adj = std::max<uint32_t>(kXcds, std::min<uint32_t>(c->NB, (uint32_t)std::max(occ, 1) * c->cus) / kXcds * kXcds);
  return 0;
}

int inu_adj_backward(inu_ctx* c, const float* seeds, float* grad_n, float* grad_z, float* grad_state) {
  int rc = check_ready(c);
  if (rc) return rc;
  auto& d = c->adj;
  if (!d.on) return fail(c, INU_ESTATE, "no adjoint recording (inu_adj_begin)");
  if (!seeds || !grad_n || !grad_z) return fail(c, INU_EINVAL, "null array");
  uint32_t K = 0;
  if ((rc = adj_recorded_steps(c, &K)) || (rc = adj_work(c))) return rc;
  const size_t NP = (size_t)c->NB * kBlockCells;
  const size_t L = c->h_ldz.size();
  HIPCHK(c, hipMemsetAsync(d.lam, 0, 3 * NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(d.gn, 0, NP * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(d.gdz, 0, std::max<size_t>(L, 1) * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(d.cur, 0, sizeof(AdjCur), c->stream));
  HIPCHK(c, hipMemcpyAsync(d.dseed, seeds, (size_t)d.max_samples * d.n_gauges * 4, hipMemcpyHostToDevice,
                           c->stream));
  d.replayed[0] = d.replayed[1] = 0ull;
  if (d.slots) {
    if ((rc = adj_backward_binomial(c, K))) return rc;
  } else {
    // Each segment from its checkpoint: replayed writing the tape, then reversed.
    for (uint32_t cs = (K + d.seg - 1) / d.seg; cs-- > 0;)
      if ((rc = adj_restore(c, cs)) || (rc = adj_turn(c, cs, K))) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  AdjCur cur;
  if ((rc = copy_sync(c, &cur, d.cur, sizeof(AdjCur), hipMemcpyDeviceToHost))) return rc;
  for (int q = 0; q < 3; ++q) d.stats[q] = cur.stats[q];
  // Manning gradient, and the state gradient, in the caller's cell order.
  rc = download_cells(c, d.gn, grad_n);
  if (!rc && grad_state)
    for (int q = 0; q < 3 && !rc; ++q) rc = download_cells(c, d.lam + q * NP, grad_state + (size_t)q * c->N);
  if (rc) return rc;
  // Bed gradient: d loss / d (dzL, dzR) of each edge, from the stream
  // position that kept it, through the face-on-the-higher-bed rule.
  std::vector<float2> hg(L);
  if (L && (rc = copy_sync(c, hg.data(), d.gdz, L * 8, hipMemcpyDeviceToHost))) return rc;
  std::vector<double> gz(NP, 0.0);
  for (size_t q = 0; q < L; ++q) {
    if (hg[q].x == 0.0f && hg[q].y == 0.0f) continue;
    const uint32_t ed = c->stream_edge[q];
    const int32_t r = c->e_cr[ed];
    if (r < 0) continue;
    const int32_t l = c->e_cl[ed];
    if (c->h_ldz[q] > 0.0f) {  // zR > zL: dzL = zR - zL
      gz[r] += hg[q].x;
      gz[l] -= hg[q].x;
    } else {  // zL >= zR: dzR = zL - zR
      gz[l] += hg[q].y;
      gz[r] -= hg[q].y;
    }
  }
  for (uint32_t s = 0; s < c->n_scen; ++s)
    for (uint32_t i = c->cell0[s]; i < c->cell0[s + 1]; ++i)
      grad_z[i] = (float)gz[(size_t)c->block0[s] * kBlockCells + (i - c->cell0[s])];
  return 0;
}

int inu_adj_replay(inu_ctx* c, float* series_out) {
  int rc = check_ready(c);
  if (rc) return rc;
  auto& d = c->adj;
  if (!d.on) return fail(c, INU_ESTATE, "no adjoint recording (inu_adj_begin)");
  if (!series_out) return fail(c, INU_EINVAL, "null output");
  uint32_t K = 0;
  if ((rc = adj_recorded_steps(c, &K)) || (rc = adj_work(c))) return rc;
  HIPCHK(c, hipMemsetAsync(d.rseries, 0, (size_t)std::max(d.max_samples, 1u) * d.n_gauges * 4, c->stream));
  if ((rc = adj_restore(c, 0)) || (rc = adj_replay_range(c, 0, K, 0, false, d.rseries))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return copy_sync(c, series_out, d.rseries, (size_t)d.max_samples * d.n_gauges * 4, hipMemcpyDeviceToHost);
}

int inu_adj_dt_log(inu_ctx* c, uint32_t* out) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->adj.on) return fail(c, INU_ESTATE, "no adjoint recording (inu_adj_begin)");
  if (!out) return fail(c, INU_EINVAL, "null output");
  return copy_sync(c, out, c->adj.log, (size_t)c->adj.max_steps * 4, hipMemcpyDeviceToHost);
}

int inu_adj_stats(inu_ctx* c, unsigned long long* out3) {
  if (!c) return INU_EINVAL;
  if (!out3) return fail(c, INU_EINVAL, "null output");
  for (int q = 0; q < 3; ++q) out3[q] = c->adj.stats[q];
  return 0;
}

int inu_adj_set_slots(inu_ctx* c, uint32_t slots) {
  if (!c) return INU_EINVAL;
  if (c->adj.on) return fail(c, INU_ESTATE, "inu_adj_set_slots: set before inu_adj_begin (a recording is open)");
  if (slots > kMaxSlots) return fail(c, INU_EINVAL, "inu_adj_set_slots: at most 64 checkpoint slots");
  c->adj_slots = slots;
  return 0;
}

int inu_adj_replayed(inu_ctx* c, unsigned long long* out2) {
  if (!c) return INU_EINVAL;
  if (!out2) return fail(c, INU_EINVAL, "null output");
  out2[0] = c->adj.replayed[0];
  out2[1] = c->adj.replayed[1];
  return 0;
}

int inu_adj_schedule_cost(uint32_t n_segments, uint32_t slots, unsigned long long* out) {
  if (!out || n_segments == 0 || slots == 0 || slots > kMaxSlots || n_segments > kMaxScheduleSegments)
    return INU_EINVAL;
  *out = revolve_schedule(n_segments, std::min(slots - 1, n_segments - 1)).at(n_segments,
                                                                             std::min(slots - 1, n_segments - 1));
  return 0;
}

int inu_adj_set_thin(inu_ctx* c, float depth) {
  if (!c) return INU_EINVAL;
  if (std::isnan(depth) || !std::isfinite(depth)) return fail(c, INU_EINVAL, "thin-layer depth must be finite");
  c->adj_thin = depth;
  for (auto it = c->graphs.begin(); it != c->graphs.end();) {
    if ((it->first & ~kKeyLenMask) == kKeyReverse) {
      (void)hipGraphExecDestroy(it->second);
      it = c->graphs.erase(it);
    } else {
      ++it;
    }
  }
  return 0;
}

int inu_adj_end(inu_ctx* c) {
  int rc = check_ready(c);
  if (rc) return rc;
  adj_release(c);
  return 0;
}

int inu_probe_vjp(inu_ctx* c, uint32_t n, const float* in17, float* out8, float g, float dry_tol) {
  if (!c) return INU_EINVAL;
  if (!in17 || !out8) return fail(c, INU_EINVAL, "null array");
  if (n == 0) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  float *din = nullptr, *dout = nullptr;
  HIPCHK(c, hipMalloc(&din, (size_t)n * 17 * 4));
  hipError_t e = hipMalloc(&dout, (size_t)n * 8 * 4);
  if (e == hipSuccess) e = hipMemcpyAsync(din, in17, (size_t)n * 17 * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    Phys pp = make_phys(g, dry_tol, 0.8f);
    pp.adj_thin = c->adj_thin < 0.0f ? kAdjThinDry * dry_tol : c->adj_thin;
    hipLaunchKernelGGL(k_probe_vjp, dim3((n + 255) / 256), dim3(256), 0, c->stream, din, dout, n, pp);
    e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out8, dout, (size_t)n * 8 * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return fail(c, INU_EHIP, std::string("probe_vjp: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
