// netpass.hip - the learned preconditioner's network pass on gfx950.
//
// Every level network of the multilevel preconditioner is C^T C with
// C = L1 L0: masked linear 3x3x3 layers, W0 lifting channel 0 to 16 channels
// and W1 mixing 16 -> 16, evaluated as the four layers (W0, W1, flip W1,
// flip W0). This file runs those four layers in three passes:
//   first:  g  = M (W1 * (M (W0 * u)))        u: single-channel float32 grid
//   middle: g2 = M (flip W1 * g)
//   last:   y  = M (flip W0 * g2)[channel 0]  y: single-channel float32 grid
// (M: the level's voxel mask; * : 3x3x3 correlation, zero padding on clamped
// axes, wrap-around on periodic ones). The first pass computes the W0 layer
// on its tile's halo in LDS from the single input channel, so that hidden
// activation never goes to memory. One launch of a pass serves every level
// of the hierarchy and every grid of a batch; a block finds its level from
// the per-level block offsets.
//
// Tiles: a block of 4 waves owns 16 x 8 x 4 (x, y, z) output voxels. Wave
// (wy, wz) computes 4 rows (y) in each of 2 planes (z) of 16 x-consecutive
// voxels; each 16-voxel row is one MFMA accumulator. The tile's input halo
// (18 x 10 x 6 voxels, all 16 channels, bf16) sits in LDS as two channel
// halves per plane [half][voxel][8 channels], so a lane's MFMA operand is one
// 16-byte read and the 16 lanes of an operand row read 256 consecutive bytes.
//
// Halo planes live in a ring of LDS plane slots. A block runs consecutive
// z-tiles of a column, and the next tile's halo shares two planes with the
// current one: those stay in their slots and only the four new planes are
// loaded (six at the start of a column). The first pass keeps its u planes
// and its hidden-layer planes in rings the same way, so W0 runs once per
// halo plane of a column, not once per tile.
//
// MFMA v_mfma_f32_16x16x32_bf16 (wave64), lane l = j + 16 q (j < 16, q < 4):
//   A (16 x 32, M x K): a[i] = A[j][8 q + i]
//   B (32 x 16, K x N): b[i] = B[8 q + i][j]
//   D (16 x 16, M x N): d[i] = D[4 q + i][j]
// (checked on the device by cc_mfma32_probe against a host product). With
// A = weights (M = output channel) and B = activations (N = voxel), lane
// j + 16 q ends with output channels 4q .. 4q+3 of voxel j: one 8-byte store,
// no transpose.
//
// K = 32 is two taps of 16 input channels. For a halo plane (tap dz) and a
// halo row Y the wave reads two operands:
//   ra[Y]: q < 2: voxel x + j of row Y, q >= 2: voxel x + j + 1 of row Y
//          -> taps (dy, dx = 0) and (dy, dx = 1) for the output row Y - dy;
//   rb[Y]: q < 2: voxel x + j + 2 of row Y, q >= 2: the same of row Y + 1
//          -> taps (dy = 0, dx = 2) and (dy = 1, dx = 2) for output row Y,
//             and (-, -) / (dy = 2, dx = 2) for output row Y - 1 (the first
//             half's weights zero).
// Every output row takes 5 MFMAs per tap plane, 15 in all (27 taps = 13.5
// MFMAs of K = 32): ra of three rows and rb of two. The 15 weight operands
// stay in registers for the whole launch.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int kFirst = 0, kMiddle = 1, kLast = 2;

constexpr int TX = 16, RW = 4, PW = 2, WY = 2, WZ = 2;
constexpr int TY = WY * RW, TZ = WZ * PW;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HP = HX * HY;       // voxels of one halo plane
constexpr int HV = HP * HZ;       // halo voxels of a tile
constexpr int NT = 64 * WY * WZ;  // threads per block
constexpr int NW = WY * WZ;       // waves per block
// The first pass stages u (float32) on the halo of the halo.
constexpr int UX = TX + 4, UY = TY + 4, UZ = TZ + 4;
constexpr int UP = UX * UY;
constexpr int UV = UP * UZ;
// Lane segments of a plane: one LDS-DMA instruction moves 64 lanes' data of
// one plane (and one channel half) into consecutive LDS slots.
constexpr int HSEG = (HP + 63) / 64;  // 3
constexpr int USEG = (UP + 63) / 64;  // 4
constexpr int NMASK = HZ * HSEG;       // mask loads of a whole halo (one byte per lane)
constexpr int MREG = (NMASK + NW - 1) / NW;
constexpr int PL16 = 2 * HP;  // 16-byte slots of one 16-channel halo plane [half][voxel]
// Ring slots. Halo plane P (0 .. HZ-1) of an item sits in slot (zc + P) mod
// the ring size, and the plane counter zc advances by TZ when the next item
// continues the column (its planes 0, 1 are the current planes TZ, TZ+1) and
// by HZ otherwise. The next item's new planes are loaded while the current
// item is computed, so they must land in slots the current item does not use:
// zc + HZ .. zc + 2 HZ - 1 at most, hence 2 HZ slots for the input ring and a
// power of two at least that for the mask ring. The hidden-layer ring of the
// first pass is written only after the block has finished the current item,
// so it needs HZ slots plus the rounding to a power of two. The u ring steps
// by TZ or UZ, with 2 UZ slots. ZMOD, a multiple of every ring size, keeps the
// counters small.
constexpr int RING = 2 * HZ;   // 12
constexpr int MRING = 16;
constexpr int HRING = 8;
constexpr int URING = 2 * UZ;  // 16
constexpr int ZMOD = 48;
static_assert(MRING >= 2 * HZ && HRING >= HZ && URING >= 2 * UZ, "ring sizes");
static_assert(ZMOD % RING == 0 && ZMOD % MRING == 0 && ZMOD % HRING == 0, "counter modulus");
// First pass: the W0 layer runs plane by plane in groups of 16 in-plane
// voxels (HG groups, the last partly past the plane); each wave takes HGW
// groups of every plane. The u ring mirrors its first UMIR slots after the
// last, so the three u planes of a tap column are linear in LDS.
constexpr int HG = (HP + 15) / 16;  // 12
constexpr int HGW = HG / NW;        // 3
constexpr int UMIR = 2;
static_assert(HG % NW == 0 && HZ % 2 == 0 && TZ % 2 == 0, "first-pass W0 plane pairs");
static_assert((MRING & (MRING - 1)) == 0 && (HRING & (HRING - 1)) == 0 && (URING & (URING - 1)) == 0,
              "power-of-two rings");

__device__ inline uint16_t to_bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Grid coordinate of a halo position: itself inside the grid, wrapped on a
// periodic axis, -1 past a clamped one. The division runs only for the few
// lanes outside the grid.
__device__ inline int wrap_index(int i, int n, int periodic) {
  if ((unsigned)i < (unsigned)n) return i;
  if (!periodic) return -1;
  i %= n;
  return i < 0 ? i + n : i;
}

// Block barrier for LDS only. __syncthreads() also releases global memory,
// which waits for every store in flight (the previous item's outputs) before
// the barrier; the passes share nothing through global memory inside a launch,
// so waiting for this wave's LDS accesses is enough. The memory clobber keeps
// the compiler from moving LDS accesses across it.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ inline bf16x8 as_bf8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ inline bf16x8 pack8(const uint16_t (&h)[8]) {
  bf16x8 v;
  __builtin_memcpy(&v, h, 16);
  return v;
}

__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace

struct NetLevel {
  int32_t nx, ny, nz, px, py, pz;
  int32_t tiles_x, tiles_y, tiles_z;
  int32_t items, block0, ipb;
  int64_t vox;            // voxels per grid (batch stride of single-channel grids)
  const uint8_t* mask;    // [vox], shared by the batch
  const uint16_t* w;      // this pass's 16 -> 16 layer, bf16 [27][16][16]
  const uint16_t* w0;     // first pass: the W0 layer (input channel 0 only)
  const void* in;         // first: float32 [nb][vox]; otherwise bf16 [nb][vox][16]
  void* out;              // last: float32 [nb][vox]; otherwise bf16 [nb][vox][16]
};

struct NetLaunch {
  NetLevel lv[CC_ML_MAX_LEVELS];
  int32_t nlev, nb;
  int32_t nblocks;  // logical blocks (the launch rounds up to a multiple of 8)
};

// Work item of a level: (grid b, column of tiles (tx, ty), z-tile) with z
// fastest, then the column (tx fastest), then the grid. A block runs a run of
// consecutive items, so its items share their z-halo planes; blocks with
// consecutive logical indices (neighbouring columns of the same grid) run on
// the same XCD and can share their x/y halos in its L2 (see logical_block).
struct ItemPos {
  int b, tz, tx, ty;
};

__device__ inline ItemPos item_decode(const NetLevel& V, int it) {
  const int tz = it % V.tiles_z;
  const int rest = it / V.tiles_z;
  const int cols = V.tiles_x * V.tiles_y;
  const int col = rest % cols;
  ItemPos P;
  P.tz = tz;
  P.b = rest / cols;
  P.ty = col / V.tiles_x;
  P.tx = col - P.ty * V.tiles_x;
  return P;
}

__device__ inline void item_next(const NetLevel& V, ItemPos& P) {
  if (++P.tz == V.tiles_z) {
    P.tz = 0;
    if (++P.tx == V.tiles_x) {
      P.tx = 0;
      if (++P.ty == V.tiles_y) {
        P.ty = 0;
        ++P.b;
      }
    }
  }
}

// Blocks are dealt to the 8 XCDs round robin (block i to XCD i mod 8); the
// logical index gives each XCD a contiguous range of logical blocks.
constexpr int kXcds = 8;
__device__ inline int logical_block(int bid, int nblocks) {
  const int per = nblocks / kXcds;  // the launch is a multiple of kXcds blocks
  return (bid % kXcds) * per + bid / kXcds;
}

__device__ inline uint32_t pack_bf16(float a, float b) {
  const bf16x2 h = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, h);
}

// s_waitcnt vmcnt(0): every global load of this wave has landed, LDS-DMA
// included.
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// LDS-DMA: lane l's BYTES bytes from base + byte_off (base the same for the
// whole wave) go to LDS slot lds_base + l * BYTES (lds_base the same for the
// whole wave; it goes to M0). Written as asm so
// that the compiler does not order the compute phase's LDS reads of the other
// buffer behind these writes: the kernel waits for them itself (wait_vmem)
// before the barrier that hands a buffer to the compute phase, and no lane
// both receives a DMA slot and writes it directly. The instruction writes LDS
// only, no registers; nothing else in these kernels uses M0.
__device__ inline uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

template <int BYTES>
__device__ inline void dma(const void* base, uint32_t byte_off, void* lds_base) {
  const uint32_t m0 = (uint32_t)__builtin_amdgcn_readfirstlane(lds_addr(lds_base));
  // readfirstlane returns int: each half goes through uint32_t so that the
  // low half is zero-extended, not sign-extended, into the 64-bit base.
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)base);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)((uintptr_t)base >> 32));
  const uint64_t sb = (uint64_t)lo | ((uint64_t)hi << 32);
  if constexpr (BYTES == 16)
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(m0), "v"(byte_off), "s"(sb) : "memory");
  else
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, %2" ::"s"(m0), "v"(byte_off), "s"(sb) : "memory");
}

template <int MODE>
__global__ __launch_bounds__(NT, 2) void k_net_pass(const NetLaunch L) {
  // Input ring: 16-channel halo planes as [slot][half][voxel] 16-byte slots.
  // The first pass instead holds float32 u planes in uring and computes its
  // hidden-layer planes into ring. Halo masks, one byte per voxel, in mring.
  __shared__ __align__(16) u32x4 ring[(MODE == kFirst ? HRING : RING) * PL16];
  __shared__ __align__(16) float uring[MODE == kFirst ? (URING + UMIR) * UP : 4];
  __shared__ __align__(16) uint8_t mring[MRING * HP];

  const int lb = logical_block((int)blockIdx.x, (int)gridDim.x);
  if (lb >= L.nblocks) return;
  int l = 0;
#pragma unroll 1
  while (l + 1 < L.nlev && lb >= L.lv[l + 1].block0) ++l;
  const NetLevel& V = L.lv[l];
  const int nx = V.nx, ny = V.ny, nz = V.nz;
  const int nxy = nx * ny;
  const int px = V.px, py = V.py, pz = V.pz;
  const int64_t vox = V.vox;
  const uint8_t* __restrict__ mask = V.mask;
  const int it0 = (lb - V.block0) * V.ipb;
  const int it1 = min(V.items, it0 + V.ipb);
  if (it0 >= it1) return;

  const int tid = threadIdx.x;
  // The wave index is the same on all lanes of a wave; readfirstlane tells
  // the compiler so, and the index arithmetic derived from it (DMA planes and
  // segments, tile rows) runs on the scalar unit.
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int j = lane & 15, q = lane >> 4;
  const int wy = wave % WY, wz = wave / WY;

  // Weight operands (A = W[tap][ci][co] with row = co = j, column k = 8q + i).
  bf16x8 WA[9], WB1[3], WB2[3];
  if constexpr (MODE != kLast) {
    const uint16_t* __restrict__ w = V.w;
    const int ci0 = 8 * (q & 1);
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const int t = dz * 9 + dy * 3 + (q >> 1);
        uint16_t h[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) h[i] = w[(t * 16 + ci0 + i) * 16 + j];
        WA[dz * 3 + dy] = pack8(h);
      }
      const int t1 = dz * 9 + (q >> 1) * 3 + 2;
      const int t2 = dz * 9 + 8;
      uint16_t h1[8], h2[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        h1[i] = w[(t1 * 16 + ci0 + i) * 16 + j];
        h2[i] = w[(t2 * 16 + ci0 + i) * 16 + j];
        if (q < 2) h2[i] = 0;
      }
      WB1[dz] = pack8(h1);
      WB2[dz] = pack8(h2);
    }
  }
  // Last pass (16 -> 1 channels): the MFMA's M dimension runs over targets
  // m = 4 dz + r (tap plane dz, output row r; m >= 12 zero) instead of output
  // channels, so one MFMA per operand covers every tap that operand feeds.
  // Row m = j of A for the ra operand of halo row Y holds W[dz][Y - r][dx][ci][0]
  // at k = 8q + i (ci = 8 (q % 2) + i, dx = q / 2), zero where Y - r is not a
  // tap row; for the rb operand (dx = 2; rows Y and Y + 1 as q / 2 = 0, 1)
  // output row r = Y takes taps dy = 0, 1 and r = Y - 1 takes dy = 2 from
  // row Y + 1, as in the other passes.
  bf16x8 LA[RW + 2], LB[RW + 1];
  if constexpr (MODE == kLast) {
    const uint16_t* __restrict__ w = V.w;
    const int tdz = j >> 2, tr = j & 3, ci0 = 8 * (q & 1), s1 = q >> 1;
#pragma unroll
    for (int Y = 0; Y < RW + 2; ++Y) {
      const int dy = Y - tr;
      const bool ok = j < 12 && dy >= 0 && dy <= 2;
      const int t = ok ? (tdz * 3 + dy) * 3 + s1 : 0;
      uint16_t h[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = ok ? w[(t * 16 + ci0 + i) * 16] : (uint16_t)0;
      LA[Y] = pack8(h);
    }
#pragma unroll
    for (int Y = 0; Y <= RW; ++Y) {
      const int dy = tr == Y ? s1 : (tr == Y - 1 && s1 == 1 ? 2 : -1);
      const bool ok = j < 12 && dy >= 0;
      const int t = ok ? (tdz * 3 + dy) * 3 + 2 : 0;
      uint16_t h[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = ok ? w[(t * 16 + ci0 + i) * 16] : (uint16_t)0;
      LB[Y] = pack8(h);
    }
  }
  // First pass: A = W0[tap = 8q + i][0][co = j] (taps past 26 zero); for
  // this lane's 8 taps the u offset from the tap column's first plane
  // (toff), and for its W0 groups the in-plane halo voxel (hrem), its u-halo
  // offset (hin) and whether it lies in the plane (hval, as a bit).
  bf16x8 W0A;
  int toff[8], hrem[HGW], hin[HGW];
  uint32_t hval = 0;
  if (MODE == kFirst) {
    uint16_t h[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int t = 8 * q + i;
      const int tt = t < 27 ? t : 26;
      h[i] = V.w0[tt * 256 + j];
      if (t >= 27) h[i] = 0;
      toff[i] = (tt / 9) * UP + ((tt / 3) % 3) * UX + tt % 3;
    }
    W0A = pack8(h);
#pragma unroll
    for (int k = 0; k < HGW; ++k) {
      const int rem = (wave + NW * k) * 16 + j;
      const int rc = rem < HP ? rem : HP - 1;
      hrem[k] = rc;
      hin[k] = (rc / HX) * UX + rc % HX;
      hval |= (uint32_t)(rem < HP) << k;
    }
  }

  // Per tile column (tx, ty) and lane: the in-plane grid offset gy * nx + gx
  // of halo plane position sg * 64 + lane (hoff) and of u-halo position
  // (uoff), with a bit per entry for "inside the plane and the grid". They
  // change only when a block moves to another column; within a column every
  // address is a per-plane scalar base plus these offsets.
  uint32_t hoff[HSEG], uoff[MODE == kFirst ? USEG : 1];
  uint32_t hok = 0, uok = 0;
  int col_tx = -1, col_ty = -1;
  auto column_tables = [&](const ItemPos& P) {
    if (P.tx == col_tx && P.ty == col_ty) return;
    col_tx = P.tx;
    col_ty = P.ty;
    const int x0 = P.tx * TX, y0 = P.ty * TY;
    hok = 0;
#pragma unroll
    for (int sg = 0; sg < HSEG; ++sg) {
      const int vp = sg * 64 + lane;
      const int gx = wrap_index(x0 - 1 + vp % HX, nx, px);
      const int gy = wrap_index(y0 - 1 + vp / HX, ny, py);
      const bool ok = vp < HP && gx >= 0 && gy >= 0;
      hoff[sg] = ok ? (uint32_t)(gy * nx + gx) : 0u;
      hok |= (uint32_t)ok << sg;
    }
    if (MODE == kFirst) {
      uok = 0;
#pragma unroll
      for (int sg = 0; sg < USEG; ++sg) {
        const int vp = sg * 64 + lane;
        const int gx = wrap_index(x0 - 2 + vp % UX, nx, px);
        const int gy = wrap_index(y0 - 2 + vp / UX, ny, py);
        const bool ok = vp < UP && gx >= 0 && gy >= 0;
        uoff[sg] = ok ? (uint32_t)(gy * nx + gx) : 0u;
        uok |= (uint32_t)ok << sg;
      }
    }
  };
  auto in_slot = [](int c) { return MODE == kFirst ? (c & (HRING - 1)) : (c % RING); };

  // Issue the loads of item P's halo planes h0 .. HZ-1 (h0 = 2 when P
  // continues the current item's column) into the ring slots of plane
  // counter zcn (u counter ucn): LDS-DMA of the input, zeros written directly
  // where the halo leaves a clamped axis, and the halo mask into registers
  // (written to LDS by put_mask).
  uint32_t mreg[MREG];
  int mk_h0 = 0, mk_zc = 0;
  auto fetch = [&](const ItemPos& P, int h0, int zcn, int ucn) {
    column_tables(P);
    const int z0 = P.tz * TZ;
    if (MODE == kFirst) {
      // The new h planes h0 .. HZ-1 read u planes h0 .. UZ-1; for a tile that
      // continues its column, u planes h0 .. TZ-1 are the current tile's
      // planes TZ+h0 .. 2 TZ-1, already in the ring.
      const int u0 = h0 ? TZ : 0;
      const float* __restrict__ src = (const float*)V.in + (size_t)P.b * vox;
#pragma unroll 1
      for (int k = wave; k < (UZ - u0) * USEG; k += NW) {
        const int uz = u0 + k / USEG, sg = k % USEG;
        const int gz = wrap_index(z0 - 2 + uz, nz, pz);
        const int us = (ucn + uz) & (URING - 1);
        float* slots = uring + us * UP + sg * 64;
        float* mirror = uring + (us + URING) * UP + sg * 64;
        if (sg * 64 + lane < UP) {
          if (gz >= 0 && ((uok >> sg) & 1)) {
            dma<4>(src + (size_t)gz * nxy, uoff[sg] * 4u, slots);
            if (us < UMIR) dma<4>(src + (size_t)gz * nxy, uoff[sg] * 4u, mirror);
          } else {
            slots[lane] = 0.f;
            if (us < UMIR) mirror[lane] = 0.f;
          }
        }
      }
    } else {
      const u32x4* __restrict__ src = (const u32x4*)V.in + (size_t)P.b * vox * 2;
#pragma unroll 1
      for (int k = wave; k < (HZ - h0) * 2 * HSEG; k += NW) {
        const int hz = h0 + k / (2 * HSEG), h = (k / HSEG) & 1, sg = k % HSEG;
        const int gz = wrap_index(z0 - 1 + hz, nz, pz);
        u32x4* slots = ring + in_slot(zcn + hz) * PL16 + h * HP + sg * 64;
        if (sg * 64 + lane < HP) {
          if (gz >= 0 && ((hok >> sg) & 1))
            dma<16>(src + 2 * (size_t)gz * nxy + h, hoff[sg] * 32u, slots);
          else
            slots[lane] = u32x4{0u, 0u, 0u, 0u};
        }
      }
    }
    mk_h0 = h0;
    mk_zc = zcn;
#pragma unroll
    for (int km = 0; km < MREG; ++km) {
      const int k = wave + NW * km;
      uint32_t m = 0;
      if (k < (HZ - h0) * HSEG) {
        const int hz = h0 + k / HSEG, sg = k % HSEG;
        const int gz = wrap_index(z0 - 1 + hz, nz, pz);
        if (gz >= 0 && ((hok >> sg) & 1)) m = (mask + (size_t)gz * nxy)[hoff[sg]];
      }
      mreg[km] = m;
    }
  };
  auto put_mask = [&]() {
#pragma unroll
    for (int km = 0; km < MREG; ++km) {
      const int k = wave + NW * km;
      if (k < (HZ - mk_h0) * HSEG) {
        const int hz = mk_h0 + k / HSEG, sg = k % HSEG;
        if (sg * 64 + lane < HP) mring[((mk_zc + hz) & (MRING - 1)) * HP + sg * 64 + lane] = (uint8_t)(mreg[km] != 0);
      }
    }
  };

  // Results of the previous item, stored after the next item's loads are
  // issued: on this target loads and stores share one completion counter, so
  // waiting for the prefetched input also waits for every store issued
  // before that wait; stores issued one compute phase earlier have landed.
  // With them: the item's first output plane, grid, and per row the in-plane
  // offset of this lane's output voxel (a bit per row inside the grid).
  uint2 pend[PW][RW];
  float pend_f[PW][RW];
  uint32_t pend_off[RW];
  uint32_t pend_ok = 0;
  int pend_z = 0, pend_b = 0;
  auto store_pending = [&]() {
#pragma unroll
    for (int p = 0; p < PW; ++p) {
      const int gz = pend_z + p;
      if (gz >= nz) continue;
      const size_t zb = (size_t)pend_b * vox + (size_t)gz * nxy;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        if (!((pend_ok >> r) & 1)) continue;
        if (MODE == kLast) {
          if (q == 0) ((float*)V.out + zb)[pend_off[r]] = pend_f[p][r];
        } else {
          reinterpret_cast<uint2*>((uint16_t*)V.out + zb * 16)[pend_off[r] * 4u + q] = pend[p][r];
        }
      }
    }
  };

  ItemPos cur = item_decode(V, it0);
  ItemPos nxt = cur;
  int zc = 0, uc = 0, h0cur = 0;
  fetch(cur, 0, zc, uc);
#pragma unroll 1
  for (int it = it0; it < it1; ++it) {
    // This item's input (issued one item ago) has landed; the block then
    // sees every wave's share of it.
    wait_vmem();
    put_mask();
    lds_barrier();
    int zn = zc, un = uc, h0n = 0;
    if (it + 1 < it1) {
      item_next(V, nxt);
      // item_next steps z first: a nonzero z-tile continues this column.
      h0n = nxt.tz != 0 ? 2 : 0;
      zn = (zc + (h0n ? TZ : HZ)) % ZMOD;
      un = (uc + (h0n ? TZ : UZ)) % ZMOD;
      fetch(nxt, h0n, zn, un);
    }
    if (it > it0) store_pending();

    if (MODE == kFirst) {
      // h = M (W0 * u) on the new halo planes h0cur .. HZ-1, two planes at a
      // time, 16 voxels per MFMA: B = the 27 taps of u (K, padded to 32) for
      // voxel j of the group. All loads of the two planes go out first.
#pragma unroll 1
      for (int hz = h0cur; hz < HZ; hz += 2) {
        float uv[2][HGW][8];
        bool live[2][HGW];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const float* ub = uring + ((uc + hz + a) & (URING - 1)) * UP;
          const uint8_t* mb = mring + ((zc + hz + a) & (MRING - 1)) * HP;
#pragma unroll
          for (int k = 0; k < HGW; ++k) {
#pragma unroll
            for (int i = 0; i < 8; ++i) uv[a][k][i] = ub[hin[k] + toff[i]];
            live[a][k] = mb[hrem[k]] != 0;
          }
        }
        f32x4 hacc[2][HGW];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int k = 0; k < HGW; ++k) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
              if (8 * q + i >= 27) uv[a][k][i] = 0.f;
            u32x4 tv;
            tv[0] = pack_bf16(uv[a][k][0], uv[a][k][1]);
            tv[1] = pack_bf16(uv[a][k][2], uv[a][k][3]);
            tv[2] = pack_bf16(uv[a][k][4], uv[a][k][5]);
            tv[3] = pack_bf16(uv[a][k][6], uv[a][k][7]);
            hacc[a][k] = mfma32(W0A, as_bf8(tv), f32x4{0.f, 0.f, 0.f, 0.f});
          }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          // channels 4q .. 4q+3: half q / 2, slot 4 (q % 2)
          uint16_t* hb = reinterpret_cast<uint16_t*>(ring + in_slot(zc + hz + a) * PL16 + (q >> 1) * HP) + 4 * (q & 1);
#pragma unroll
          for (int k = 0; k < HGW; ++k) {
            if (!((hval >> k) & 1)) continue;
            uint2 raw;
            raw.x = live[a][k] ? pack_bf16(hacc[a][k][0], hacc[a][k][1]) : 0u;
            raw.y = live[a][k] ? pack_bf16(hacc[a][k][2], hacc[a][k][3]) : 0u;
            *reinterpret_cast<uint2*>(hb + 8 * hrem[k]) = raw;
          }
        }
      }
      lds_barrier();
    }

    f32x4 acc[PW][RW];
    f32x4 lacc[PW + 2];  // last pass: D[4 dz + r][x] of each halo plane
#pragma unroll
    for (int p = 0; p < PW; ++p)
#pragma unroll
      for (int r = 0; r < RW; ++r) acc[p][r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int hz = 0; hz < PW + 2; ++hz) lacc[hz] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Per halo plane the wave's operands (ra of RW + 2 rows, rb of RW + 1)
    // are read together, then consumed by the plane's MFMAs: with two waves
    // per SIMD, a read issued just before its MFMA leaves the LDS latency
    // exposed.
#pragma unroll
    for (int hz = 0; hz < PW + 2; ++hz) {
      const int slot = __builtin_amdgcn_readfirstlane(in_slot(zc + wz * PW + hz));
      const u32x4* pl = ring + slot * PL16 + (q & 1) * HP;
      bf16x8 ra[RW + 2], rb[RW + 1];
#pragma unroll
      for (int Y = 0; Y < RW + 2; ++Y) {
        const int row = (wy * RW + Y) * HX;
        ra[Y] = as_bf8(pl[row + j + (q >> 1)]);
        if (Y <= RW) rb[Y] = as_bf8(pl[row + (q >> 1) * HX + j + 2]);
      }
      // Keeps the scheduler from sinking the reads next to their MFMAs.
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (MODE == kLast) {
#pragma unroll
        for (int Y = 0; Y < RW + 2; ++Y) {
          lacc[hz] = mfma32(LA[Y], ra[Y], lacc[hz]);
          if (Y <= RW) lacc[hz] = mfma32(LB[Y], rb[Y], lacc[hz]);
        }
        continue;
      }
#pragma unroll
      for (int Y = 0; Y < RW + 2; ++Y) {
#pragma unroll
        for (int p = 0; p < PW; ++p) {
          const int dz = hz - p;
          if (dz < 0 || dz > 2) continue;
#pragma unroll
          for (int r = 0; r < RW; ++r) {
            const int dy = Y - r;
            if (dy >= 0 && dy <= 2) acc[p][r] = mfma32(WA[dz * 3 + dy], ra[Y], acc[p][r]);
          }
          if (Y < RW) acc[p][Y] = mfma32(WB1[dz], rb[Y], acc[p][Y]);
          if (Y >= 1 && Y <= RW) acc[p][Y - 1] = mfma32(WB2[dz], rb[Y], acc[p][Y - 1]);
        }
      }
    }

    // Masked, rounded results stay in registers until the next item's loads
    // are issued.
    pend_z = cur.tz * TZ + wz * PW;
    pend_b = cur.b;
    pend_ok = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int gx = cur.tx * TX + j, gy = cur.ty * TY + wy * RW + r;
      const bool ok = gx < nx && gy < ny;
      pend_off[r] = ok ? (uint32_t)(gy * nx + gx) : 0u;
      pend_ok |= (uint32_t)ok << r;
    }
#pragma unroll
    for (int p = 0; p < PW; ++p) {
      const uint8_t* mp = mring + ((zc + wz * PW + p + 1) & (MRING - 1)) * HP;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const bool live = mp[(wy * RW + r + 1) * HX + j + 1] != 0;
        if (MODE == kLast) {
          // Output plane p, row r of voxel j: lane (j, q) holds tap plane
          // dz = q of halo plane p + q; the three are summed across lanes.
          float v = 0.f;
          if (q == 0) v = lacc[p][r];
          if (q == 1) v = lacc[p + 1][r];
          if (q == 2) v = lacc[p + 2][r];
          v += __shfl_xor(v, 16, 64);
          v += __shfl_xor(v, 32, 64);
          pend_f[p][r] = live ? v : 0.f;
        } else {
          pend[p][r].x = live ? pack_bf16(acc[p][r][0], acc[p][r][1]) : 0u;
          pend[p][r].y = live ? pack_bf16(acc[p][r][2], acc[p][r][3]) : 0u;
        }
      }
    }
    cur = nxt;
    zc = zn;
    uc = un;
    h0cur = h0n;
  }
  store_pending();
}

// D = A B for bf16 A (16 x 32) and B (32 x 16), row-major, fp32 D (16 x 16):
// the 16x16x32 MFMA lane layout the passes rely on.
__global__ void k_mfma32_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, float* __restrict__ D) {
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, q = lane >> 4;
  uint16_t a[8], b[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = A[j * 32 + 8 * q + i];
    b[i] = B[(8 * q + i) * 16 + j];
  }
  const f32x4 d = mfma32(pack8(a), pack8(b), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(4 * q + i) * 16 + j] = d[i];
}

// Blocks per pass launch: the items of every level are split into runs of the
// same length, at most g_net_blocks runs in all levels together
// (cc_net_set_config; by default two per CU of the current device, the most
// that fit by LDS), so that every block is resident from the start. A block
// beyond that starts only when another ends and doubles the launch time.
static int g_net_blocks = 0;
static int g_net_legacy = 0;
constexpr int kBlocksPerCu = 2;
constexpr int kMaxDevices = 64;
static int g_resident[kMaxDevices];

static int resident_blocks() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return kBlocksPerCu * 256;
  if (g_resident[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    g_resident[dev] = kBlocksPerCu * cus;
  }
  return g_resident[dev];
}

bool net_pass_legacy() { return g_net_legacy != 0; }

// The three passes on the levels l with has_net[l]: u[l] -> g1[l] -> g2[l] ->
// y[l], each holding nb grids one after the other.
int net_pass(const cc_mlcnn_desc& ml, float* const* u, uint16_t* const* g1, uint16_t* const* g2, float* const* y,
             int nb, hipStream_t s) {
  if (nb < 1) return CC_E_ARG;
  NetLaunch L[3];
  for (int m = 0; m < 3; ++m) {
    L[m] = NetLaunch{};
    L[m].nb = nb;
  }
  long long total = 0;
  int nlev = 0;
  int lev_of[CC_ML_MAX_LEVELS];
  for (int l = 0; l < ml.levels; ++l) {
    if (!ml.has_net[l]) continue;
    const cc_cnn_desc& d = ml.net[l];
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0 || !d.mask) return CC_E_ARG;
    for (int k = 0; k < 4; ++k)
      if (!d.weights[k]) return CC_E_ARG;
    NetLevel V{};
    V.nx = d.nx;
    V.ny = d.ny;
    V.nz = d.nz;
    V.px = d.periodic[0];
    V.py = d.periodic[1];
    V.pz = d.periodic[2];
    V.tiles_x = (d.nx + TX - 1) / TX;
    V.tiles_y = (d.ny + TY - 1) / TY;
    V.tiles_z = (d.nz + TZ - 1) / TZ;
    const long long items = (long long)V.tiles_x * V.tiles_y * V.tiles_z * nb;
    if (items > (1LL << 30)) return CC_E_ARG;
    V.items = (int)items;
    V.vox = (int64_t)d.nx * d.ny * d.nz;
    V.mask = d.mask;
    total += items;
    lev_of[nlev] = l;
    for (int m = 0; m < 3; ++m) L[m].lv[nlev] = V;
    ++nlev;
  }
  if (nlev == 0) return 0;
  // The shortest run length whose per-level block counts (each rounded up)
  // fit in max_blocks; one block per level when the levels outnumber them.
  const long long max_blocks = g_net_blocks > 0 ? g_net_blocks : resident_blocks();
  long long max_items = 1;
  for (int k = 0; k < nlev; ++k) max_items = L[0].lv[k].items > max_items ? L[0].lv[k].items : max_items;
  long long ipb_all = (total + max_blocks - 1) / max_blocks;
  if (ipb_all < 1) ipb_all = 1;
  for (;;) {
    long long nblk = 0;
    for (int k = 0; k < nlev; ++k) nblk += (L[0].lv[k].items + ipb_all - 1) / ipb_all;
    if (nblk <= max_blocks || ipb_all >= max_items) break;
    ++ipb_all;
  }
  int block0 = 0;
  for (int k = 0; k < nlev; ++k) {
    const int l = lev_of[k];
    const cc_cnn_desc& d = ml.net[l];
    const long long ipb = ipb_all;
    const int blocks = (int)((L[0].lv[k].items + ipb - 1) / ipb);
    for (int m = 0; m < 3; ++m) {
      NetLevel& V = L[m].lv[k];
      V.block0 = block0;
      V.ipb = (int)ipb;
    }
    L[0].lv[k].w = d.weights[1];
    L[0].lv[k].w0 = d.weights[0];
    L[0].lv[k].in = u[l];
    L[0].lv[k].out = g1[l];
    L[1].lv[k].w = d.weights[2];
    L[1].lv[k].in = g1[l];
    L[1].lv[k].out = g2[l];
    L[2].lv[k].w = d.weights[3];
    L[2].lv[k].in = g2[l];
    L[2].lv[k].out = y[l];
    block0 += blocks;
  }
  const int launch = (block0 + kXcds - 1) / kXcds * kXcds;
  for (int m = 0; m < 3; ++m) {
    L[m].nlev = nlev;
    L[m].nblocks = block0;
  }
  hipLaunchKernelGGL(k_net_pass<kFirst>, dim3((unsigned)launch), dim3(NT), 0, s, L[0]);
  hipLaunchKernelGGL(k_net_pass<kMiddle>, dim3((unsigned)launch), dim3(NT), 0, s, L[1]);
  hipLaunchKernelGGL(k_net_pass<kLast>, dim3((unsigned)launch), dim3(NT), 0, s, L[2]);
  return (int)hipGetLastError();
}

}  // namespace cc

using namespace cc;

extern "C" {

int cc_mfma32_probe(const uint16_t* A, const uint16_t* B, float* D, cc_stream_t stream) {
  if (!A || !B || !D) return CC_E_ARG;
  hipLaunchKernelGGL(k_mfma32_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
  return (int)hipGetLastError();
}

int cc_net_set_config(int32_t blocks, int32_t legacy) {
  if (blocks < 0 || blocks > (1 << 20) || (legacy != 0 && legacy != 1)) return CC_E_ARG;
  g_net_blocks = blocks;
  g_net_legacy = legacy;
  return 0;
}

size_t cc_net_level_workspace_bytes(const cc_cnn_desc* net, int32_t nb) {
  if (!net || nb < 1) return 0;
  const size_t v = (size_t)net->nx * net->ny * net->nz * (size_t)nb;
  return 2 * (((v * 16 * sizeof(uint16_t)) + 255) & ~size_t(255));
}

int cc_net_level_forward(const cc_cnn_desc* net, int32_t nb, const float* u, float* y, void* workspace,
                         size_t workspace_bytes, cc_stream_t stream) {
  if (!net || !u || !y || !workspace || nb < 1) return CC_E_ARG;
  if (workspace_bytes < cc_net_level_workspace_bytes(net, nb)) return CC_E_WORKSPACE;
  const size_t v = (size_t)net->nx * net->ny * net->nz * (size_t)nb;
  const size_t half = ((v * 16 * sizeof(uint16_t)) + 255) & ~size_t(255);
  cc_mlcnn_desc ml{};
  ml.levels = 1;
  ml.has_net[0] = 1;
  ml.net[0] = *net;
  float* uu[CC_ML_MAX_LEVELS] = {const_cast<float*>(u)};
  uint16_t* g1[CC_ML_MAX_LEVELS] = {(uint16_t*)workspace};
  uint16_t* g2[CC_ML_MAX_LEVELS] = {(uint16_t*)((char*)workspace + half)};
  float* yy[CC_ML_MAX_LEVELS] = {y};
  return net_pass(ml, uu, g1, g2, yy, nb, (hipStream_t)stream);
}

}  // extern "C"
