/*
 * fr_apf.hip — potential-field intrinsic costs of a graph (src/cuda/apf.cu).
 *
 * Repulsion (apf.cu:23-98): every obstacle pixel outside the ego box sets
 * its own intrinsic cost to 100*Kr/2; an obstacle pixel with a free
 * 4-neighbour (an edge pixel) then adds Kr/2*(1/p - 1/p0)^2 to every other
 * pixel outside the ego box within Euclidean distance p <= p0 = radius.
 * The reference scatters with float atomics from one thread per source, so
 * its sums are unordered and its set races with the other threads' adds.
 * Here the result is fixed: every set comes first and each pixel adds its
 * contributions in ascending source-pixel order, which is one outcome of
 * the reference and exactly the CPU oracle's order.
 *
 * MI355X shape: a gather, one lane per target pixel, instead of the
 * reference's scatter of (2r+1)^2 atomics per source.
 *   k_apf_sources  edge pixels as a bit plane, one 64-bit word per wave
 *                  (ballot over 64 pixels of a row);
 *   k_apf_table    the contribution of each offset (|dx|, |dz|), and per
 *                  |dz| the largest |dx| inside the disc (the disc rows are
 *                  intervals because sqrtf is monotone);
 *   k_apf_gather   per target row of the disc, the words of the edge plane
 *                  are walked set bit by set bit in ascending x, so the
 *                  float adds run in ascending source order.
 * Attraction (apf.cu:100-121) is elementwise.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "fr_engine.h"
#include "fr_math.h"

namespace {

__device__ __forceinline__ bool apf_in_ego(const FrApfArgs &a, int x, int z)
{
    return x >= a.lbx && x <= a.ubx && z >= a.ubz && z <= a.lbz;
}

/* classCosts[TO_INT(og.x)] < 0 (apf.cu:17-21,44-46); a class id outside the
 * table counts as an obstacle, as in the planner's frame preparation. */
__device__ __forceinline__ bool apf_obstacle(const FrApfArgs &a, long pos)
{
    const int cls = fr_trunc_f(a.xyz[3 * pos]);
    return cls < 0 || cls >= a.n_classes || a.costs[cls] < 0;
}

/* apf.cu:17-21: outside the grid counts as an obstacle */
__device__ __forceinline__ bool apf_obstacle_xz(const FrApfArgs &a, int x, int z)
{
    if (x < 0 || x >= a.W || z < 0 || z >= a.H)
        return true;
    return apf_obstacle(a, (long)z * a.W + x);
}

} // namespace

/* Edge pixels (apf.cu:35-53): obstacle, outside the ego box, and not
 * enclosed by obstacles on all four sides.  One wave per 64-pixel word. */
__global__ void __launch_bounds__(256) k_apf_sources(FrApfArgs a, uint64_t *__restrict__ bits, int wpr)
{
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= (long)a.H * wpr)
        return;
    const int z = (int)(wave / wpr), wi = (int)(wave - (long)z * wpr);
    const int x = wi * 64 + lane;
    bool src = false;
    if (x < a.W && !apf_in_ego(a, x, z) && apf_obstacle(a, (long)z * a.W + x))
        src = !(apf_obstacle_xz(a, x - 1, z) && apf_obstacle_xz(a, x + 1, z) && apf_obstacle_xz(a, x, z - 1) &&
                apf_obstacle_xz(a, x, z + 1));
    const unsigned long long m = __ballot(src);
    if (lane == 0)
        bits[wave] = m;
}

/* Contribution of offset (dx, dz), 0 <= dx, dz <= rr, in the reference's
 * operation order (apf.cu:84-88), and per dz the largest dx with p <= p0
 * (-1 if none).  The self offset is never read. */
__global__ void __launch_bounds__(256) k_apf_table(float kr_half, int radius, int rr, float *__restrict__ tab,
                                                   int *__restrict__ span)
{
    const float p0 = (float)radius;
    const long n = (long)(rr + 1) * (rr + 1);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int dz = (int)(i / (rr + 1)), dx = (int)(i - (long)dz * (rr + 1));
        const float p = sqrtf((float)(dx * dx + dz * dz));
        float c = 0.0f;
        if (!(p > p0) && (dx | dz)) {
            const float f = 1 / p - 1 / p0;
            c = kr_half * f * f;
        }
        tab[i] = c;
        if (dx == 0) {
            int m = -1;
            for (int t = 0; t <= rr; ++t) {
                if (sqrtf((float)(t * t + dz * dz)) > p0)
                    break;
                m = t;
            }
            span[dz] = m;
        }
    }
}

/* One lane per target pixel, 64 pixels of a row per wave. */
__global__ void __launch_bounds__(256) k_apf_gather(FrApfArgs a, float kr_half, int rr,
                                                    const uint64_t *__restrict__ bits, int wpr,
                                                    const float *__restrict__ tab, const int *__restrict__ span)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int z = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || z >= a.H || apf_in_ego(a, x, z))
        return;
    const long pos = (long)z * a.W + x;
    float acc = apf_obstacle(a, pos) ? 100 * kr_half : a.intrinsic[pos];
    const int z0 = z - rr < 0 ? 0 : z - rr, z1 = z + rr >= a.H ? a.H - 1 : z + rr;
    for (int sz = z0; sz <= z1; ++sz) {
        const int dz = sz > z ? sz - z : z - sz;
        const int m = span[dz];
        if (m < 0)
            continue;
        const int x0 = x - m < 0 ? 0 : x - m, x1 = x + m >= a.W ? a.W - 1 : x + m;
        const uint64_t *row = bits + (long)sz * wpr;
        const float *trow = tab + (long)dz * (rr + 1);
        for (int wi = x0 >> 6; wi <= (x1 >> 6); ++wi) {
            uint64_t w = row[wi];
            const int lo = x0 - wi * 64, hi = x1 - wi * 64;
            if (lo > 0)
                w &= ~0ull << lo;
            if (hi < 63)
                w &= (2ull << hi) - 1ull;
            while (w) {
                const int sx = wi * 64 + __builtin_ctzll(w);
                w &= w - 1ull;
                if (sx == x && dz == 0)
                    continue;
                acc += trow[sx > x ? sx - x : x - sx];
            }
        }
    }
    a.intrinsic[pos] = acc;
}

/* apf.cu:100-121: free pixels lose Ka/2 * squared distance to the goal. */
__global__ void __launch_bounds__(256) k_apf_attraction(FrApfArgs a, float ka_half, int gx, int gz)
{
    const long n = (long)a.W * a.H;
    for (long pos = (long)blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += (long)gridDim.x * blockDim.x) {
        if (apf_obstacle(a, pos))
            continue;
        const int z = (int)(pos / a.W), x = (int)(pos - (long)z * a.W);
        const int dx = gx - x, dz = gz - z;
        const float d = (float)(dx * dx + dz * dz) * ka_half;
        a.intrinsic[pos] = a.intrinsic[pos] - d;
    }
}

int fr_apf_window(int radius, int W, int H)
{
    if (radius < 1)
        return 0;
    const int lim = (W > H ? W : H) - 1;
    return radius < lim ? radius : lim;
}

void fr_launch_apf_repulsion(const FrApfArgs &a, float kr_half, int radius, uint64_t *d_bits, float *d_tab,
                             int *d_span, hipStream_t st)
{
    const int wpr = (a.W + 63) / 64;
    const long words = (long)a.H * wpr;
    const int rr = fr_apf_window(radius, a.W, a.H);
    hipLaunchKernelGGL(k_apf_sources, dim3((unsigned)((words + 3) / 4)), dim3(256), 0, st, a, d_bits, wpr);
    const long nt = (long)(rr + 1) * (rr + 1);
    const long tb = (nt + 255) / 256;
    hipLaunchKernelGGL(k_apf_table, dim3((unsigned)(tb < 4096 ? tb : 4096)), dim3(256), 0, st, kr_half, radius, rr,
                       d_tab, d_span);
    hipLaunchKernelGGL(k_apf_gather, dim3((unsigned)wpr, (unsigned)((a.H + 3) / 4)), dim3(256), 0, st, a, kr_half,
                       rr, (const uint64_t *)d_bits, wpr, (const float *)d_tab, (const int *)d_span);
}

void fr_launch_apf_attraction(const FrApfArgs &a, float ka_half, int gx, int gz, hipStream_t st)
{
    const long n = (long)a.W * a.H;
    const long b = (n + 255) / 256;
    hipLaunchKernelGGL(k_apf_attraction, dim3((unsigned)(b < 8192 ? b : 8192)), dim3(256), 0, st, a, ka_half, gx,
                       gz);
}

/* ------------------------------------------------------------------------ */
/* Allocation-free repulsion: the same three kernels, their scratch (edge   */
/* bit plane, offset table, disc spans) carved from one file-scope device   */
/* array instead of hipMalloc'd buffers.                                    */
/* ------------------------------------------------------------------------ */

/* Each GPU holds its own copy of this array until the process ends.  Host
 * code never passes its name to a kernel: the name is a host-side stand-in,
 * the device address comes from hipGetSymbolAddress (fr_apf_scratch_ptr). */
static __device__ __align__(256) unsigned char scratch[FR_APF_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lookup_mu;
void *g_scratch_ptr[kMaxDevices];
/* calls on one device share its array: they are serialised here, and a
 * call returns only after its kernels have finished */
std::mutex g_scratch_use_mu[kMaxDevices];

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

} // namespace

size_t fr_apf_scratch_bytes(int W, int H, int radius)
{
    const size_t words = (size_t)H * (size_t)((W + 63) / 64);
    const size_t w = (size_t)fr_apf_window(radius, W, H);
    return round256(8 * words) + round256(4 * (w + 1) * (w + 1)) + round256(4 * (w + 1));
}

/* Device address of `scratch` on the current device, looked up once per
 * device (the first call on it) with that device current. */
static unsigned char *fr_apf_scratch_ptr(int *dev_out)
{
    int dev = -1;
    FR_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("compute_apf_repulsion_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_lookup_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("compute_apf_repulsion_noalloc: hipGetSymbolAddress failed: ") +
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *dev_out = dev;
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

void fr_launch_apf_repulsion_noalloc(const FrApfArgs &a, float kr_half, int radius, hipStream_t st)
{
    const size_t need = fr_apf_scratch_bytes(a.W, a.H, radius);
    if (need > FR_APF_SCRATCH_BYTES)
        throw std::runtime_error("compute_apf_repulsion_noalloc: the edge plane and offset table need " +
                                 std::to_string(need) + " bytes of scratch, more than its " +
                                 std::to_string((size_t)FR_APF_SCRATCH_BYTES) +
                                 " (use compute_apf_repulsion for this grid and radius)");
    int dev = -1;
    unsigned char *base = fr_apf_scratch_ptr(&dev);
    const size_t words = (size_t)a.H * (size_t)((a.W + 63) / 64);
    const size_t w = (size_t)fr_apf_window(radius, a.W, a.H);
    uint64_t *bits = reinterpret_cast<uint64_t *>(base);
    float *tab = reinterpret_cast<float *>(base + round256(8 * words));
    int *span = reinterpret_cast<int *>(base + round256(8 * words) + round256(4 * (w + 1) * (w + 1)));
    /* every word, table entry and span the gather reads is written by the
     * two kernels before it, so nothing left by an earlier call is read */
    std::lock_guard<std::mutex> lk(g_scratch_use_mu[dev]);
    fr_launch_apf_repulsion(a, kr_half, radius, bits, tab, span, st);
    FR_CHECK(hipGetLastError());
    FR_CHECK(hipStreamSynchronize(st));
}
