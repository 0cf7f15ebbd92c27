// fx_kern32.h -- kernel arguments of the FP32-mixed path, shared by the
// single-domain launches and the batched ones.
#pragma once

#include "fx_params.h"

// pointers of one domain's FP32-mixed state as the step sweep sees them
struct s32p {
    const double *d_in[2];
    const float2 *q_in[2];          // (qx, qy)
    const double *zb;               // bed elevation
    float *f6;
    float2 *ff;                     // (fe1, fn1) of the step
    uint8_t *flags;
    float2 *tw;                     // wet-time tracer (closed runs, open-run start)
    double *partials;
    int64_t mm;
};

// pointers of one domain's state as the fast ADE sweep sees them
struct a32p {
    const double *d[2];
    const float2 *ff;
    const uint8_t *flags;
    float *mass[2][FX_MAX_CHEM];
    const float *melt_mult;
    int64_t mm;
    int k0;                   // first solute of this launch
};

// The fold of a single domain's CFL partials into FX32_REC records by the
// first FX32_REC blocks of the fast ADE launch that follows the step sweep,
// for the next iteration's clock (rec == NULL: no fold)
struct a32fold {
    const double *partials;
    double *rec;
    int32_t npart, nrec;
};
#define FX32_REC 64

// wave reductions of one double (64 lanes, butterfly: every lane gets the result)
__device__ inline double wmin(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// tiling of the FP32 kernels
// The FP32 kernels address a cell's element as a uniform base plus a 32-bit
// byte offset (cell index times element size), which the hardware takes as a
// scalar base and a 32-bit vector offset.  Domains are limited to
// FX32_MAX_CELLS so that the offset of an 8-byte element fits.
#define FX32_MAX_CELLS ((int64_t)1 << 29)
// The access goes through a global-memory pointer: every array the FP32
// kernels address this way is device memory, and a pointer read from memory
// (a batch member's slot) would otherwise be a generic one, whose loads count
// on the LDS counter as well, so that every LDS wait would also wait for them.
// The host pass of these inline device functions sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define FX_GLOBAL __attribute__((address_space(1)))
#else
#define FX_GLOBAL
#endif
template <typename T>
__device__ inline FX_GLOBAL T &at_off(T *base, uint32_t off)
{
    return *(FX_GLOBAL T *)(reinterpret_cast<char *>(base) + off);
}
template <typename T>
__device__ inline const FX_GLOBAL T &at_off(const T *base, uint32_t off)
{
    return *(const FX_GLOBAL T *)(reinterpret_cast<const char *>(base) + off);
}

// finiteness by the exponent bits, which no floating-point option of the
// compiler can fold away (the compiler emits them as one class test)
__device__ inline bool fx_finite(double v)
{
    return (((unsigned long long)__double_as_longlong(v) >> 52) & 0x7ffull) != 0x7ffull;
}
__device__ inline bool fx_finitef(float v)
{
    return (((unsigned)__float_as_uint(v) >> 23) & 0xffu) != 0xffu;
}

// One-instruction forms for the FP32 kernels.  max(0, x), and min / max of
// two floats of which at least one is non-negative (>= +0), compare the bit
// patterns as signed integers: for such pairs the integer order is the float
// order, and no NaN quieting of the operands is needed as for fmaxf / fminf.
// The reciprocal, square root and reciprocal square root are the hardware's
// (1 ulp; FP32 denormals are flushed in these kernels anyway).
__device__ inline float fx_max0(float x) { return __int_as_float(max(0, __float_as_int(x))); }
__device__ inline float fx_max_nn(float a, float b) { return __int_as_float(max(__float_as_int(a), __float_as_int(b))); }
__device__ inline float fx_min_nn(float a, float b) { return __int_as_float(min(__float_as_int(a), __float_as_int(b))); }
__device__ inline float fx_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float fx_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// a uniform 64-bit lane mask as this lane's flag: no instruction, the mask is
// used as the lane predicate (the inverse of a ballot)
__device__ inline bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

#define FX32_TH 61            // rows owned per step-sweep strip
#define FX32_CH 32            // default columns per step-sweep block
#define FX32_CH_MIN 8         // smallest step-sweep block (sizes the CFL partials)
#define FX32_AOWN 58          // rows owned per ADE-sweep strip
#define FX32_ACH 64           // default columns per ADE-sweep block
#define FX32_FOLD_MAX 256     // blocks of the clock's partial fold
#define FX32_NPV 4            // values per step-sweep partial: max wave speed, max h, weir volume, solute melt area

// One member of a batch, in device memory: everything its kernels need.
struct fx_bslot {
    fx_params P;
    s32p S;
    a32p A;
    fx_ctl *ctl;
    int32_t gx;               // step-sweep blocks along rows
    int32_t agx;              // ADE-sweep blocks along rows
    int32_t npart;            // step-sweep blocks (CFL partials)
    int32_t ade_on;
};

// Compact per-member status written after every batched step (read by the
// host to find members that finished or reached their print time).
struct fx_bstat {
    double tim, dtfl;
    int32_t it, steps, solver_steps, halt, status, pad;
};

s32p make_s32p(fx_domain d);
a32p make_a32p(fx_domain d);
int fx32_npart(fx_domain d);
int fx32_ade_gx(fx_domain d);
int fx32_ade_blocks(fx_domain d);
int fx32_set_tiling(fx_domain d, int ch, int ach);
// sums over the step-sweep partials not yet folded by the clock: weir-face
// volume [m3] and the melt area that brought solute [cells, times the multiplier]
int fx32_pending(fx_domain d, double *weir_vol, double *solute_area);

// batched launches over nslot members; tmap / amap give, per block, the
// member (high 12 bits) and the member's own block index (low 20 bits)
int fx32b_launch_clock(hipStream_t st, const fx_bslot *slots, int nslot);
int fx32b_launch_sweep(hipStream_t st, const fx_bslot *slots, const uint32_t *tmap, int nblocks);
int fx32b_launch_ade(hipStream_t st, const fx_bslot *slots, const uint32_t *amap, int nblocks, int nchem);
int fx32b_launch_end(hipStream_t st, const fx_bslot *slots, int nslot, fx_bstat *stat);

// Tile of workgroup w of an n-tile launch.  Workgroups go to the eight XCDs
// round-robin (w % 8), and each XCD has its own L2.  With chunk > 0 the tiles
// are dealt to the XCDs in runs of `chunk` consecutive tiles (strips fastest),
// round-robin: a tile's column-block neighbour, whose halo columns it
// re-reads, is usually in the same run, so on the same XCD and processed
// shortly before.  The tail that does not fill a whole round goes in one
// contiguous run per XCD.  A bijection of [0, n) for every n and chunk.
__host__ __device__ inline int fx_xcd_tile(int w, int n, int chunk)
{
    if (chunk <= 0) return w;
    const int round = 8 * chunk;
    const int nfull = (n / round) * round;
    if (w >= nfull) {
        const int m = n - nfull, v = w - nfull;
        const int x = v & 7, q = m >> 3, r = m & 7;
        return nfull + x * q + (x < r ? x : r) + (v >> 3);
    }
    const int x = w & 7, l = w >> 3;                 // l-th workgroup of XCD x
    return ((l / chunk) * 8 + x) * chunk + l % chunk;
}
// run length of the tile deal of a single domain's sweep with gx strips per
// column block and n tiles (FLUXOS_XCD_COLS = column blocks per run, default
// 4; 0 = no remap; negative = one run per XCD), and of a batch's tile maps
// (FLUXOS_XCD_BATCH = tiles per run; unset or 0 = no remap)
int fx32_xcd_chunk(int gx, int n);
int fx32_xcd_chunk_batch();

#define FX_BMAP(slot, blk) (((uint32_t)(slot) << 20) | (uint32_t)(blk))
#define FX_BMAP_SLOT(v) ((v) >> 20)
#define FX_BMAP_BLK(v) ((v) & 0xFFFFFu)
