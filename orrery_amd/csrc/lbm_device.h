// Device-side definitions shared by the D3Q19 and D2Q9 kernels: population storage formats,
// the in-place (A-A) addressing, launch-shape helpers and the lattice descriptors.
//
// In-place A-A scheme (Bailey, Myre, Walsh, Lilja and Pickett 2009). One array holds the 19 (9)
// populations of every cell, A[i][x] below (the D3Q19 layout in blocks of cells, Lattice3; the
// D2Q9 one in planes of `stride`, Lattice2). Let n be the number of completed steps.
//  - n even ("post-A" layout): the post-collision population f*_i(x) sits at A[opp(i)][x].
//  - n odd  ("post-B" layout): f*_i(x) sits at A[i][x + c_i] (in-domain fluid targets only).
// Step n even (B step): node x reads its pull values f*_i(x - c_i) at A[opp(i)][x - c_i] and
// writes f*_i(x) to A[i][x + c_i]. Step n odd (A step): node x reads A[i][x] and writes
// A[opp(i)][x]. In both, the addresses a node reads are exactly the ones it writes, so one
// array suffices and no two threads touch the same element.
//
// Boundary nodes (class 2: domain faces and neighbours of solids) also keep their 19
// post-collision values in a double-buffered stash, [2][Q][bstride], written every step
// (buffer (n+1)&1) and read in the next (buffer n&1). Every boundary rule whose value is not
// the node's own natural pull -- bounce-back, the interpolated wall, moving walls, slip
// mirrors and the outlet -- reads the stash, so in-place overwrites never race with them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_lattice.h"

namespace lbm {

// ---- storage formats ----------------------------------------------------------------------
// Each format: S the stored type, R the arithmetic type, `shifted` whether the arithmetic works
// on f - w. The binary32 and binary64 formats store with pack(v, dither) (the dither unused).
// `offset` (the 16-bit formats): the format stores a population's offset from the free-stream
// equilibrium rather than f - w itself (Lattice3::shift, Lattice2::shift); the D3Q19 kernels
// round it with 32 random bits by pack(v, scaled, dither), or two cells into one word by
// pack_pair, the D2Q9 kernels to nearest by nearest(v, scaled).
//
// The 16-bit formats round stochastically: the bits below the stored significand are raised by
// the dither and then cut, an unbiased rounding. Round-to-nearest makes the rounding error a
// function of the value alone, so every node of a uniform region gains the same error each
// step, a coherent momentum source that builds into a spurious body force of the order of the
// drag on a 33 M-node channel (bf16: 70% drag error at step 2000); with a dither that differs
// from node to node and from step to step the errors are independent and unbiased.

// Dither of a slot: lowbias32 (C. Wellons' integer hash, 2018).
__device__ __forceinline__ std::uint32_t slot_dither(std::uint32_t x) {
    x ^= x >> 16U;
    x *= 0x7feb352dU;
    x ^= x >> 15U;
    x *= 0x846ca68bU;
    x ^= x >> 16U;
    return x;
}

// 32 random bits of node `cell` in the step whose epoch (slot_dither of the lattice's step
// clock) is `epoch`: the dither pattern changes every step, so the rounding errors of one slot
// are independent from step to step as well as from node to node.
__device__ __forceinline__ std::uint32_t cell_dither(std::uint32_t cell, std::uint32_t epoch) {
    return slot_dither(cell ^ epoch);
}

// The dither of direction i from its node's bits: the node's word rotated by 7 i (19 distinct
// rotations), so each direction reads a uniformly distributed window of it; `T::kDitherRotation`
// more for a format whose conversion reads its random bits from the top of the word.
template <class T>
__device__ __forceinline__ std::uint32_t direction_dither(std::uint32_t bits, int i) {
    const int r = (7 * i + T::kDitherRotation) & 31;
    return r == 0 ? bits : (bits << r) | (bits >> (32 - r));
}

struct StoreF32 {
    using S = float;
    using R = float;
    static constexpr bool shifted = true;
    static constexpr bool dithered = false;
    static constexpr bool offset = false;
    static constexpr int id = 0;
    static constexpr int kDitherRotation = 0;
    static __device__ __forceinline__ R load(S v) { return v; }
    static __device__ __forceinline__ S pack(R v, std::uint32_t) { return v; }
    static __device__ __forceinline__ S nearest(R v) { return v; }
    static __device__ __forceinline__ S from_double(double v) { return static_cast<float>(v); }
};

// binary16 holding (f - w) * 2^15: the scale moves the normal range of binary16 down to
// [1.9e-9, 2.0], which covers every shifted population of a weakly compressible flow.
struct StoreF16 {
    using S = _Float16;
    using R = float;
    static constexpr bool shifted = true;
    static constexpr bool dithered = true;
    static constexpr bool offset = true;
    static constexpr int id = 1;
    static constexpr float kScale = 32768.0f;
    // The conversion takes its 13 random bits from bits 19-31 of the word: rotated 19 further,
    // those are bits 0-12 of the direction's 7 i rotation.
    static constexpr int kDitherRotation = 19;
    static __device__ __forceinline__ R load(S v) { return static_cast<float>(v) * (1.0f / kScale); }
    static __device__ __forceinline__ S nearest(R v) { return static_cast<_Float16>(v * kScale); }
    // Round to nearest of the value less an offset, `scaled` = offset 2^15 (the D2Q9 lattices).
    static __device__ __forceinline__ S nearest(R v, R scaled) {
        return static_cast<_Float16>(__builtin_fmaf(v, kScale, -scaled));
    }
    static __device__ __forceinline__ S from_double(double v) {
        return static_cast<_Float16>(static_cast<float>(v * 32768.0));
    }
    // With an offset: v = s / 2^15 + offset, one fused multiply-add (the power-of-two scale is
    // exact, so this is the rounding of the plain sum); stored: (v - offset) 2^15 as
    // v 2^15 - offset 2^15, `scaled` = offset 2^15, the same rounding.
    static __device__ __forceinline__ R load(S v, R offset) {
        return __builtin_fmaf(static_cast<float>(v), 1.0f / kScale, offset);
    }
    // Stochastic rounding by the hardware conversion (v_cvt_sr_f16_f32): in binary16's normal
    // range it adds bits 19-31 of the random word to the 13 binary32 significand bits below
    // binary16's and cuts them, bit for bit (u + (word >> 19)) & ~0x1FFF as an integer on the
    // binary32 pattern u, measured over 2.8 M values on gfx950; below that range it rounds at the
    // subnormal spacing, also stochastically.
    static __device__ __forceinline__ S pack(R v, R scaled, std::uint32_t dither) {
        using Pair = __fp16 __attribute__((ext_vector_type(2)));
        const R x = __builtin_fmaf(v, kScale, -scaled);
        const Pair h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(Pair, x), x, dither, false);
        return __builtin_bit_cast(S, h[0]);
    }
    // Two values rounded into the low and the high half of one word (the first conversion keeps
    // the high half of the word it writes, here its own source, which the second replaces).
    static __device__ __forceinline__ std::uint32_t pack_pair(R v0, R v1, R scaled, std::uint32_t d0, std::uint32_t d1) {
        using Pair = __fp16 __attribute__((ext_vector_type(2)));
        const R x0 = __builtin_fmaf(v0, kScale, -scaled);
        Pair h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(Pair, x0), x0, d0, false);
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, __builtin_fmaf(v1, kScale, -scaled), d1, true);
        return __builtin_bit_cast(std::uint32_t, h);
    }
    static constexpr float kOffsetScale = kScale;
};

// bfloat16 holding f - w, rounded to nearest even.
struct StoreBF16 {
    using S = std::uint16_t;
    using R = float;
    static constexpr bool shifted = true;
    static constexpr bool dithered = true;
    static constexpr bool offset = true;
    static constexpr int id = 2;
    static constexpr int kDitherRotation = 0;
    static __device__ __forceinline__ R load(S v) {
        return __uint_as_float(static_cast<std::uint32_t>(v) << 16U);
    }
    static __device__ __forceinline__ S nearest(R v) {
        const std::uint32_t u = __float_as_uint(v);
        return static_cast<S>((u + 0x7FFFU + ((u >> 16U) & 1U)) >> 16U);
    }
    // Round to nearest even of the value less an offset (`scaled` is the offset itself; the
    // D2Q9 lattices).
    static __device__ __forceinline__ S nearest(R v, R scaled) { return nearest(v - scaled); }
    // Initial states round to nearest even.
    static __device__ __forceinline__ S from_double(double v) { return nearest(static_cast<float>(v)); }
    // With an offset (no scale: `scaled` is the offset itself).
    static __device__ __forceinline__ R load(S v, R offset) { return load(v) + offset; }
    // Stochastic rounding by the hardware conversion (v_cvt_sr_bf16_f32): it adds the top 16
    // bits of the random word to the 16 binary32 significand bits below bfloat16's and cuts
    // them, bit for bit (u + (word >> 16)) >> 16 on the binary32 pattern u, measured over 4.2 M
    // values on gfx950.
    static __device__ __forceinline__ S pack(R v, R scaled, std::uint32_t dither) {
        using Pair = __bf16 __attribute__((ext_vector_type(2)));
        const R x = v - scaled;
        const Pair b = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(Pair, x), x, dither, false);
        return __builtin_bit_cast(S, b[0]);
    }
    // Two values rounded into the low and the high half of one word (the first conversion keeps
    // the high half of the word it writes, here its own source, which the second replaces).
    static __device__ __forceinline__ std::uint32_t pack_pair(R v0, R v1, R scaled, std::uint32_t d0, std::uint32_t d1) {
        using Pair = __bf16 __attribute__((ext_vector_type(2)));
        const R x0 = v0 - scaled;
        Pair b = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(Pair, x0), x0, d0, false);
        b = __builtin_amdgcn_cvt_sr_bf16_f32(b, v1 - scaled, d1, true);
        return __builtin_bit_cast(std::uint32_t, b);
    }
    static constexpr float kOffsetScale = 1.0f;
};

// binary64, unshifted: the on-device oracle, with the reference CPU solver's arithmetic.
struct StoreF64 {
    using S = double;
    using R = double;
    static constexpr bool shifted = false;
    static constexpr bool dithered = false;
    static constexpr bool offset = false;
    static constexpr int id = 3;
    static constexpr int kDitherRotation = 0;
    static __device__ __forceinline__ R load(S v) { return v; }
    static __device__ __forceinline__ S pack(R v, std::uint32_t) { return v; }
    static __device__ __forceinline__ S nearest(R v) { return v; }
    static __device__ __forceinline__ S from_double(double v) { return v; }
};

// ---- node classes ----------------------------------------------------------------------------
// 0 solid; 1 bulk (all 18 neighbours in the domain and fluid); 2 boundary (listed: faces and
// neighbours of solids, updated by the boundary kernel through source codes and the stash);
// 3 side (a fluid node on a y or z face, away from the x faces and from solids, whose every
// source lies in the lattice after the side's mirror or wrap: streamed by the bulk kernel in A
// steps and the side kernel in B steps); 4 inlet (a node of the velocity inlet face x = 0, off
// the y and z faces and away from solids: streamed by the bulk kernel like a bulk node, the
// regularised inlet rule replacing the pulls from x = -1). The reference's classification
// (classify_nodes) counts side and inlet nodes as boundary nodes.
enum NodeClass : std::uint8_t { kSolidNode = 0, kBulkNode = 1, kListedNode = 2, kSideNode = 3, kInletNode = 4 };

// ---- boundary-node source codes -------------------------------------------------------------
// One 32-bit word per (boundary node, direction), built on the host from resolve_d3q19:
//   bits 0..2 kind (SourceKind), bit 3 mirror_y applied, bit 4 mirror_z applied,
//   bits 5..31 payload: kNode -> source cell (a bulk or side node, read from the lattice);
//   kStash / kOutlet -> source boundary slot (read from the stash); kWall -> side (0..3).
// The D2Q9 tables (resolve_d2q9) have no side nodes: there a mirrored kNode carries the source
// boundary slot and is read from the stash.
constexpr std::uint32_t kCodeMirrorY = 8U;
constexpr std::uint32_t kCodeMirrorZ = 16U;
constexpr int kCodeShift = 5;

__host__ __device__ constexpr std::uint32_t make_code(int kind, bool my, bool mz,
                                                      std::uint32_t payload) {
    return static_cast<std::uint32_t>(kind) | (my ? kCodeMirrorY : 0U) | (mz ? kCodeMirrorZ : 0U) |
           (payload << kCodeShift);
}

// ---- lattice descriptor (kernel argument) ---------------------------------------------------
template <typename R> struct Lattice3 {
    // Populations in blocks of 2^block_shift consecutive cells, each block [19][2^block_shift]:
    // A[i][c] is element (c' >> k) 19 2^k + i 2^k + (c' mod 2^k), c' = c + guard, k = block_shift.
    // `guard` (whole blocks) cells before the first cell and at least as many after the last
    // keep the plain addresses of every cell's neighbours (nx ny + nx + 1 cells away, and a
    // run's V - 1 more) inside the buffer. A wave's 19 accesses of an A step then fall in one
    // span of 19 2^k elements instead of 19 planes across the buffer.
    void* f;
    const std::uint8_t* cls;  // node class per cell (NodeClass)
    const std::uint32_t* bcell;  // boundary node -> cell [nb]
    const std::uint32_t* bcode;  // [19][bstride]
    const std::uint32_t* bwrite; // [bstride] bit i: the lattice copy of output i has a reader
    const R* bq;                 // [19][bstride] wall fractions (interpolated wall) or null
    R* stash;                    // [2][19][bstride]
    R* hist;                     // [5][bstride] convective-outlet history (outlet nodes)
    const std::uint32_t* tile_slot;  // [cells / 64 + 1] boundary nodes before each 64-cell run
    const std::uint32_t* scell;      // side node -> cell [ns] (cell order): the side kernel's list
    const std::uint32_t* clock;      // steps since initialisation (the 16-bit formats' dither epoch)
    std::uint32_t nx, ny, nz, cells, nb, bstride, ns;
    std::uint32_t block_shift;  // k: 2^k cells per block
    std::uint32_t guard;        // guard cells before the first cell (a multiple of 2^k)
    FastDiv div_nx, div_plane;
    int x_mode;       // 0 periodic, 1 inflow/outflow
    int convective;   // outlet blends with history
    int side[4];      // y lo, y hi, z lo, z hi wall codes
    R wall_u[4];      // wall speeds along x
    R inlet[3];       // imposed inlet velocity
    R tau;            // relaxation time
    R omega;          // 1 / tau (closure off)
    R smag_c;         // Smagorinsky constant (CPU form: K c c)
    R smag_c2;        // its square (shifted form: K c^2)
    int les;
    R force[3];
    int forced;
    R outlet_speed;
    R sponge_strength;
    std::uint32_t sponge_in, sponge_out, sponge_sides;
    int sponge_y, sponge_z;
    // Formats with `offset` store f_i - f_eq_i(1, u_ref): shift[i] = f_eq_i(1, u_ref) - w_i,
    // with u_ref the configured inlet velocity (zero on a periodic lattice), is added to a loaded
    // value and taken from a stored one (shift[19 + i]: the same in the stored units, scaled as the
    // format scales). The uniform stream then stores values near zero, where a 16-bit format's
    // relative precision is a small absolute error. A device table [38], not descriptor fields:
    // a lane-dependent index into the descriptor would copy it to scratch memory.
    const R* shift;
};

template <typename R> struct Lattice2 {
    void* f;                     // S[9][stride]
    const std::uint8_t* cls;
    const std::uint32_t* bcell;
    const std::uint32_t* bcode;  // [9][bstride]
    R* stash;                    // [2][9][bstride]
    R* hist;                     // [3][bstride]
    std::uint32_t nx, ny, cells, stride, nb, bstride;
    FastDiv div_nx;
    int x_mode;
    int convective;
    int side[2];
    R wall_u[2];
    R inlet[2];
    R tau, omega, smag_c, smag_c2;
    int les;
    R force[2];
    int forced;
    R outlet_speed;
    const R* shift;  // formats with `offset`: [18] each direction's offset, then scaled as stored
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (block b runs on XCD
// b % 8). Renumbering them so that each XCD owns one contiguous run of tiles keeps the cache
// lines shared by neighbouring tiles inside one XCD's L2.
__device__ __forceinline__ std::uint32_t xcd_tile(std::uint32_t b, std::uint32_t nblocks) {
    const std::uint32_t q = nblocks >> 3U;
    const std::uint32_t r = nblocks & 7U;
    const std::uint32_t xcd = b & 7U;
    const std::uint32_t local = b >> 3U;
    return xcd * q + (xcd < r ? xcd : r) + local;
}

// Largest |u| of one block (exact: a maximum does not depend on the order it is taken in).
__device__ __forceinline__ void block_max(double v, double* partial) {
    __shared__ double lds[16];
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        v = fmax(v, __shfl_down(v, offset, 64));
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        const int waves = static_cast<int>(blockDim.x >> 6);
        for (int w = 0; w < waves; ++w) {
            m = fmax(m, lds[w]);
        }
        partial[blockIdx.x] = m;
    }
}

} // namespace lbm
