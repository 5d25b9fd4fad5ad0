// mega_pre.hip -- the one-time geometry preprocessing of the model on the
// device, behind mega_preprocess (include/mega.h):
//   GET_ADIAL              PRE_MODEL_ADIAL.F90:16-158   nodal window = the
//                          largest per-axis extent of the attached elements
//   GET_VOLUMES            PRE_MODEL_VOLUMES.F90:16-121 nodal volume = 1/8 of
//                          each attached hexahedron (GET_VOL2, three
//                          determinants) or 1/4 of each tetrahedron (GET_VOL1)
//   GET_SMOOTHING_LENGTHS  PRE_MODEL_SMOOTHING_LENGTHS.F90:16-102, and the
//                          global XDIST_MAX / YDIST_MAX / ZDIST_MAX
// called in this order by PRE_MODEL.F90:248-262.
//
// The reference loops over elements and scatters into the nodes.  Here:
//   k_pre_elem   one lane per element: extents, volume share, corner counts
//   (scan)       node -> element incidence rows
//   k_pre_fill   corners appended to their node's row
//   k_pre_node   one lane per node: its row put in (element, corner) order --
//                the order of the reference's scatter -- then the max of the
//                extents, the volume as the same left-to-right sum, and the
//                smoothing lengths; per-block maxima for XDIST_MAX etc.
// Windows and volumes are bit-identical to the reference's.  VOL**THRD is
// the correctly rounded power (dd::pow_dd); the host C library's pow, which
// the reference calls, misrounds about 0.1 % of such powers by one ulp, so a
// smoothing length can differ from the reference's in its last bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <mutex>
#include <vector>

#include "ddmath.hpp"
#include "mega.h"

namespace {

#define PRECHK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "mega_pre: %s failed: %s (%s:%d)\n", #expr,  \
                         hipGetErrorString(e_), __FILE__, __LINE__);          \
            status = MEGA_ERR_HIP;                                            \
            goto done;                                                        \
        }                                                                     \
    } while (0)

constexpr int kPreBlock = 256;

// DETERMINANT (DETERMINANT.F90:28-30) of a column-major 3x3, products and
// sums rounded one by one as the host computes them (no fused multiply-add)
__device__ __forceinline__ double det3_abs_unfused(const double *a)
{
#pragma clang fp contract(off)
#define A_(i, j) a[((i)-1) + 3 * ((j)-1)]
    const double d = A_(1,1)*(A_(2,2)*A_(3,3) - A_(3,2)*A_(2,3))
                   + A_(1,2)*(A_(3,1)*A_(2,3) - A_(2,1)*A_(3,3))
                   + A_(1,3)*(A_(2,1)*A_(3,2) - A_(3,1)*A_(2,2));
#undef A_
    return d < 0.0 ? -d : d;
}

// corners of an element: 8 for a hexahedron (block type 2), the first 4 of
// the 8 ids for a tetrahedron (PRE_MODEL_ADIAL.F90:77,112)
__device__ __forceinline__ int corners(const int32_t *hex, int e) { return (!hex || hex[e]) ? 8 : 4; }

__global__ void k_pre_elem(int numel, int nump, const double *coo, const int32_t *elcon, const int32_t *hex,
                           double *ext, double *share, int32_t *count, int32_t *bad)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= numel) return;
    const int nc = corners(hex, e);
    double X[8][3];
    for (int k = 0; k < nc; ++k) {
        const int id = elcon[8 * (int64_t)e + k];
        if (id < 1 || id > nump) {   // ids are 1-based node numbers
            atomicOr(bad, 1);
            return;
        }
        X[k][0] = coo[3 * (int64_t)(id - 1)];
        X[k][1] = coo[3 * (int64_t)(id - 1) + 1];
        X[k][2] = coo[3 * (int64_t)(id - 1) + 2];
    }
    // largest |x_k - x_j| over the corner pairs (PRE_MODEL_ADIAL.F90:77-91);
    // rounding is monotone, so it is fl(max - min)
    for (int a = 0; a < 3; ++a) {
        double lo = X[0][a], hi = X[0][a];
        for (int k = 1; k < nc; ++k) {
            lo = X[k][a] < lo ? X[k][a] : lo;
            hi = X[k][a] > hi ? X[k][a] : hi;
        }
        ext[3 * (int64_t)e + a] = hi - lo;
    }
    double dvol;
    if (nc == 8) {   // GET_VOL2 (PRE_MODEL_VOL2.F90:36-52): columns 7-1 | 2-1, 5-1, 4-1 | 3-6, 6-8, 8-3
        double A[9], B[9], C[9];
        for (int a = 0; a < 3; ++a) {
            const double d71 = X[6][a] - X[0][a];
            A[a] = d71; B[a] = d71; C[a] = d71;
            A[3 + a] = X[1][a] - X[0][a];
            B[3 + a] = X[4][a] - X[0][a];
            C[3 + a] = X[3][a] - X[0][a];
            A[6 + a] = X[2][a] - X[5][a];
            B[6 + a] = X[5][a] - X[7][a];
            C[6 + a] = X[7][a] - X[2][a];
        }
        dvol = (det3_abs_unfused(A) + det3_abs_unfused(B) + det3_abs_unfused(C)) / 6.0;
        share[e] = 0.125 * dvol;
    } else {         // GET_VOL1 (PRE_MODEL_VOL1.f90): rows 4-3, 1-3, 2-3
        double A[9];
        for (int a = 0; a < 3; ++a) {
            A[0 + 3 * a] = X[3][a] - X[2][a];
            A[1 + 3 * a] = X[0][a] - X[2][a];
            A[2 + 3 * a] = X[1][a] - X[2][a];
        }
        dvol = det3_abs_unfused(A) / 6.0;
        share[e] = 0.25 * dvol;
    }
    for (int k = 0; k < nc; ++k) atomicAdd(&count[elcon[8 * (int64_t)e + k] - 1], 1);
}

// corner k of element e appended to its node's row (in arbitrary order)
__global__ void k_pre_fill(int numel, const int32_t *elcon, const int32_t *hex, int32_t *cursor, int32_t *inc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= numel) return;
    const int nc = corners(hex, e);
    for (int k = 0; k < nc; ++k) {
        const int slot = atomicAdd(&cursor[elcon[8 * (int64_t)e + k] - 1], 1);
        inc[slot] = 8 * e + k;
    }
}

__global__ void k_pre_node(int nump, const int32_t *start, int32_t *inc, const double *ext, const double *share,
                           double *win, double *vol, double *sm_len, double *dpart)
{
#pragma clang fp contract(off)
    __shared__ double s_max[3][kPreBlock];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    if (i < nump) {
        const int s = start[i], L = start[i + 1] - s;
        // (element, corner) order: the order in which the reference's element loop visits this node
        for (int a = s + 1; a < s + L; ++a) {
            const int v = inc[a];
            int b = a - 1;
            while (b >= s && inc[b] > v) { inc[b + 1] = inc[b]; --b; }
            inc[b + 1] = v;
        }
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, V = 0.0;   // ADIAL = 0, VOLUME = 0 (PRE_MODEL_ADIAL.F90:59)
        for (int a = s; a < s + L; ++a) {
            const int e = inc[a] >> 3;
            const double x0 = ext[3 * (int64_t)e], x1 = ext[3 * (int64_t)e + 1], x2 = ext[3 * (int64_t)e + 2];
            if (x0 > w0) w0 = x0;
            if (x1 > w1) w1 = x1;
            if (x2 > w2) w2 = x2;
            V = V + share[e];
        }
        win[3 * (int64_t)i] = w0; win[3 * (int64_t)i + 1] = w1; win[3 * (int64_t)i + 2] = w2;
        vol[i] = V;
        // PRE_MODEL_SMOOTHING_LENGTHS.F90:52-72
        const double thrd = 0.333333333333333;
        const double dim1 = w0 * 2.0, dim2 = w1 * 2.0, dim3 = w2 * 2.0;
        const double vol_win = dim1 * dim2 * dim3;
        const double ratio = (dd::pow_dd(V, thrd) / dd::pow_dd(vol_win, thrd)) * 1.0;
        const double l1 = dim1 * ratio / 2.0, l3 = dim2 * ratio / 2.0, l5 = dim3 * ratio / 2.0;
        if (sm_len) {
            double *o = sm_len + 6 * (int64_t)i;
            o[0] = l1; o[1] = l1; o[2] = l3; o[3] = l3; o[4] = l5; o[5] = l5;
        }
        // IF (SM_LEN(k,I) .GT. XDIST_MAX) ... (:94-96): a NaN never raises the maximum
        if (l1 > d0) d0 = l1;
        if (l3 > d1) d1 = l3;
        if (l5 > d2) d2 = l5;
    }
    s_max[0][threadIdx.x] = d0; s_max[1][threadIdx.x] = d1; s_max[2][threadIdx.x] = d2;
    __syncthreads();
    for (int h = kPreBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int a = 0; a < 3; ++a) {
                const double o = s_max[a][threadIdx.x + h];
                if (o > s_max[a][threadIdx.x]) s_max[a][threadIdx.x] = o;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) dpart[3 * blockIdx.x + a] = s_max[a][0];
}

template <class T>
int dalloc_pre(T **p, size_t count)
{
    if (count == 0) count = 1;
    return hipMalloc((void **)p, sizeof(T) * count) == hipSuccess ? MEGA_OK : MEGA_ERR_NOMEM;
}

}  // namespace

extern "C" int mega_preprocess(int32_t nump, const double *coo, int32_t numel, const int32_t *elcon,
                               const int32_t *hex, double *win, double *vol, double *sm_len, double *dist_max)
{
    if (nump <= 0 || numel < 0 || numel > INT32_MAX / 8 || !coo || (numel > 0 && !elcon) || !win || !vol ||
        !dist_max)
        return MEGA_ERR_ARG;
    int status = MEGA_OK;
    hipStream_t st = nullptr;
    double *d_coo = nullptr, *d_ext = nullptr, *d_share = nullptr, *d_win = nullptr, *d_vol = nullptr;
    double *d_sm = nullptr, *d_dpart = nullptr;
    int32_t *d_el = nullptr, *d_hex = nullptr, *d_count = nullptr, *d_start = nullptr, *d_cursor = nullptr;
    int32_t *d_inc = nullptr, *d_bad = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    const int nbe = (numel + kPreBlock - 1) / kPreBlock, nbn = (nump + kPreBlock - 1) / kPreBlock;
    int32_t bad = 0, ninc = 0;
    std::vector<double> part;
    PRECHK(hipStreamCreate(&st));
    if (dalloc_pre(&d_coo, 3 * (size_t)nump) || dalloc_pre(&d_el, 8 * (size_t)numel) ||
        dalloc_pre(&d_ext, 3 * (size_t)numel) || dalloc_pre(&d_share, (size_t)numel) ||
        dalloc_pre(&d_win, 3 * (size_t)nump) || dalloc_pre(&d_vol, (size_t)nump) ||
        (sm_len && dalloc_pre(&d_sm, 6 * (size_t)nump)) || dalloc_pre(&d_dpart, 3 * (size_t)nbn) ||
        dalloc_pre(&d_count, (size_t)nump + 1) || dalloc_pre(&d_start, (size_t)nump + 1) ||
        dalloc_pre(&d_cursor, (size_t)nump + 1) || dalloc_pre(&d_inc, 8 * (size_t)numel) ||
        dalloc_pre(&d_bad, 1) || (hex && dalloc_pre(&d_hex, (size_t)numel))) {
        status = MEGA_ERR_NOMEM;
        goto done;
    }
    PRECHK(hipMemcpyAsync(d_coo, coo, sizeof(double) * 3 * (size_t)nump, hipMemcpyHostToDevice, st));
    if (numel > 0)
        PRECHK(hipMemcpyAsync(d_el, elcon, sizeof(int32_t) * 8 * (size_t)numel, hipMemcpyHostToDevice, st));
    if (hex && numel > 0)
        PRECHK(hipMemcpyAsync(d_hex, hex, sizeof(int32_t) * (size_t)numel, hipMemcpyHostToDevice, st));
    PRECHK(hipMemsetAsync(d_count, 0, sizeof(int32_t) * ((size_t)nump + 1), st));
    PRECHK(hipMemsetAsync(d_bad, 0, sizeof(int32_t), st));
    if (numel > 0) {
        k_pre_elem<<<nbe, kPreBlock, 0, st>>>(numel, nump, d_coo, d_el, d_hex, d_ext, d_share, d_count, d_bad);
        PRECHK(hipGetLastError());
    }
    PRECHK(hipMemcpyAsync(&bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PRECHK(hipStreamSynchronize(st));
    if (bad) {
        status = MEGA_ERR_ARG;   // an element names a node that does not exist
        goto done;
    }
    PRECHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_count, d_start, nump + 1, st));
    PRECHK(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1));
    PRECHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, d_count, d_start, nump + 1, st));
    PRECHK(hipMemcpyAsync(d_cursor, d_start, sizeof(int32_t) * (size_t)nump, hipMemcpyDeviceToDevice, st));
    PRECHK(hipMemcpyAsync(&ninc, d_start + nump, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (numel > 0) {
        k_pre_fill<<<nbe, kPreBlock, 0, st>>>(numel, d_el, d_hex, d_cursor, d_inc);
        PRECHK(hipGetLastError());
    }
    k_pre_node<<<nbn, kPreBlock, 0, st>>>(nump, d_start, d_inc, d_ext, d_share, d_win, d_vol, d_sm, d_dpart);
    PRECHK(hipGetLastError());
    part.resize(3 * (size_t)nbn);
    PRECHK(hipMemcpyAsync(win, d_win, sizeof(double) * 3 * (size_t)nump, hipMemcpyDeviceToHost, st));
    PRECHK(hipMemcpyAsync(vol, d_vol, sizeof(double) * (size_t)nump, hipMemcpyDeviceToHost, st));
    if (sm_len) PRECHK(hipMemcpyAsync(sm_len, d_sm, sizeof(double) * 6 * (size_t)nump, hipMemcpyDeviceToHost, st));
    PRECHK(hipMemcpyAsync(part.data(), d_dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    PRECHK(hipStreamSynchronize(st));
    dist_max[0] = dist_max[1] = dist_max[2] = 0.0;   // XDIST_MAX = 0 (:45-47)
    for (int b = 0; b < nbn; ++b)
        for (int a = 0; a < 3; ++a)
            if (part[3 * b + a] > dist_max[a]) dist_max[a] = part[3 * b + a];
done:
    for (void *p : {(void *)d_coo, (void *)d_ext, (void *)d_share, (void *)d_win, (void *)d_vol, (void *)d_sm,
                    (void *)d_dpart, (void *)d_el, (void *)d_hex, (void *)d_count, (void *)d_start,
                    (void *)d_cursor, (void *)d_inc, (void *)d_bad, tmp})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return status;
}

// ---------------------------------------------------------------------------
// mega_preprocess_noalloc: the same kernels with every intermediate buffer in
// one file-scope device array (no hipMalloc / hipFree, no stream creation).
//
// The array lives on each GPU the library's code object is loaded on, for the
// life of the process.  Host code never passes its name to a kernel: the
// name is a host-side stand-in, not a device address.  The device address is
// taken once per device with hipGetSymbolAddress (device current) and kept in
// pre_scratch_ptr[device].  Every call on one GPU shares that GPU's array, so
// calls that overlap on one GPU race; what must start at zero (the corner
// counts and the bad-id flag) is cleared by a fill kernel at each call.
namespace {

constexpr int64_t kNoAllocMaxNodes = MEGA_PRE_NOALLOC_MAX_NODES;
constexpr int64_t kNoAllocMaxElems = MEGA_PRE_NOALLOC_MAX_ELEMS;
constexpr int64_t kNoAllocCubBytes = 256 * 1024;   // hipcub scan temporary storage
constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kNoAllocBlocks = (kNoAllocMaxNodes + kPreBlock - 1) / kPreBlock;
// ext[3 numel] + share[numel] (doubles), inc[8 numel], count/start/cursor[nump + 1],
// bad, dpart[3 blocks], dmax[3], cub temporary storage
constexpr int64_t kNoAllocBytes =
    align256(8 * 3 * kNoAllocMaxElems) + align256(8 * kNoAllocMaxElems) + align256(4 * 8 * kNoAllocMaxElems) +
    3 * align256(4 * (kNoAllocMaxNodes + 1)) + align256(4) + align256(8 * 3 * kNoAllocBlocks) + align256(8 * 3) +
    kNoAllocCubBytes;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
void *pre_scratch_ptr[kMaxDevices];
std::mutex pre_scratch_mu;

// device address of this device's scratch array (looked up once per device)
int pre_scratch(void **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return MEGA_ERR_HIP;
    std::lock_guard<std::mutex> lk(pre_scratch_mu);
    if (!pre_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "mega_preprocess_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
            return MEGA_ERR_HIP;
        }
        pre_scratch_ptr[dev] = p;
    }
    *out = pre_scratch_ptr[dev];
    return MEGA_OK;
}

__global__ void k_fill_i32(int64_t n, int32_t *p, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// XDIST_MAX, YDIST_MAX, ZDIST_MAX from the per-block maxima, starting from 0
// as PRE_MODEL_SMOOTHING_LENGTHS.F90:45-47 does; one block
__global__ void k_pre_dmax(int nblocks, const double *dpart, double *dmax)
{
    __shared__ double s_max[3][kPreBlock];
    double d[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += kPreBlock)
        for (int a = 0; a < 3; ++a)
            if (dpart[3 * b + a] > d[a]) d[a] = dpart[3 * b + a];
    for (int a = 0; a < 3; ++a) s_max[a][threadIdx.x] = d[a];
    __syncthreads();
    for (int h = kPreBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int a = 0; a < 3; ++a)
                if (s_max[a][threadIdx.x + h] > s_max[a][threadIdx.x]) s_max[a][threadIdx.x] = s_max[a][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) dmax[a] = s_max[a][0];
}

}  // namespace

extern "C" int mega_preprocess_noalloc(int32_t nump, const double *coo, int32_t numel, const int32_t *elcon,
                                       const int32_t *hex, double *win, double *vol, double *sm_len,
                                       double *dist_max)
{
    if (nump <= 0 || numel < 0 || !coo || (numel > 0 && !elcon) || !win || !vol || !dist_max) return MEGA_ERR_ARG;
    if (nump > kNoAllocMaxNodes || numel > kNoAllocMaxElems) {
        std::fprintf(stderr,
                     "mega_preprocess_noalloc: %d nodes / %d elements exceed the scratch maximum of %lld / %lld "
                     "(use mega_preprocess)\n",
                     nump, numel, (long long)kNoAllocMaxNodes, (long long)kNoAllocMaxElems);
        return MEGA_ERR_SCRATCH;
    }
    int status = MEGA_OK;
    void *base = nullptr;
    if ((status = pre_scratch(&base)) != MEGA_OK) return status;
    const hipStream_t st = nullptr;   // the device's null stream
    unsigned char *q = static_cast<unsigned char *>(base);
    auto take = [&q](int64_t bytes) { unsigned char *p = q; q += align256(bytes); return (void *)p; };
    double *d_ext = (double *)take(8 * 3 * kNoAllocMaxElems);
    double *d_share = (double *)take(8 * kNoAllocMaxElems);
    int32_t *d_inc = (int32_t *)take(4 * 8 * kNoAllocMaxElems);
    int32_t *d_count = (int32_t *)take(4 * (kNoAllocMaxNodes + 1));
    int32_t *d_start = (int32_t *)take(4 * (kNoAllocMaxNodes + 1));
    int32_t *d_cursor = (int32_t *)take(4 * (kNoAllocMaxNodes + 1));
    int32_t *d_bad = (int32_t *)take(4);
    double *d_dpart = (double *)take(8 * 3 * kNoAllocBlocks);
    double *d_dmax = (double *)take(8 * 3);
    void *d_tmp = take(kNoAllocCubBytes);
    size_t tmp_bytes = 0;
    const int nbe = (numel + kPreBlock - 1) / kPreBlock, nbn = (nump + kPreBlock - 1) / kPreBlock;
    int32_t bad = 0;
    k_fill_i32<<<(unsigned)((nump + 1 + kPreBlock - 1) / kPreBlock), kPreBlock, 0, st>>>(nump + 1, d_count, 0);
    k_fill_i32<<<1, kPreBlock, 0, st>>>(1, d_bad, 0);
    PRECHK(hipGetLastError());
    if (numel > 0) {
        k_pre_elem<<<nbe, kPreBlock, 0, st>>>(numel, nump, coo, elcon, hex, d_ext, d_share, d_count, d_bad);
        PRECHK(hipGetLastError());
    }
    PRECHK(hipMemcpyAsync(&bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PRECHK(hipStreamSynchronize(st));
    if (bad) {
        status = MEGA_ERR_ARG;   // an element names a node that does not exist
        goto done;
    }
    PRECHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_count, d_start, nump + 1, st));
    if ((int64_t)tmp_bytes > kNoAllocCubBytes) {
        std::fprintf(stderr, "mega_preprocess_noalloc: scan storage %zu B exceeds the reserved %lld B\n", tmp_bytes,
                     (long long)kNoAllocCubBytes);
        status = MEGA_ERR_SCRATCH;
        goto done;
    }
    PRECHK(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_count, d_start, nump + 1, st));
    PRECHK(hipMemcpyAsync(d_cursor, d_start, sizeof(int32_t) * (size_t)nump, hipMemcpyDeviceToDevice, st));
    if (numel > 0) {
        k_pre_fill<<<nbe, kPreBlock, 0, st>>>(numel, elcon, hex, d_cursor, d_inc);
        PRECHK(hipGetLastError());
    }
    k_pre_node<<<nbn, kPreBlock, 0, st>>>(nump, d_start, d_inc, d_ext, d_share, win, vol, sm_len, d_dpart);
    PRECHK(hipGetLastError());
    k_pre_dmax<<<1, kPreBlock, 0, st>>>(nbn, d_dpart, d_dmax);
    PRECHK(hipGetLastError());
    PRECHK(hipMemcpyAsync(dist_max, d_dmax, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    PRECHK(hipStreamSynchronize(st));
done:
    return status;
}
