/*
 * fr_frame.hip — frame preprocessing on the device.
 *
 * The reference takes its frames from libdriveless, whose SearchFrame fills
 * the distance-to-goal channel (pixel channel 1) on processDistanceToGoal()
 * (called by unittests/src/test_graph_can_connect_hermite.cpp:51 before
 * canConnectToGoal, which reads that channel as the Hermite curve's length,
 * graph_hermite_connection.cu:20).  libdriveless is not part of the
 * reference tree; here the channel is the Euclidean pixel
 * distance to the goal, sqrtf of the integer squared distance, which is the
 * length the loop's own goal connection computes when the channel is not
 * processed (graph_direct_goal_connection.cu:29-36).
 *
 * The safe-zone bit 0x100 of channel 2, which the goal connection reads at a
 * curve's start cell to skip the curve's footprint tests
 * (graph_direct_goal_connection.cu:124-135, with isSafeZoneChecked = true at
 * fastrrt.cpp:145), also comes from libdriveless.  Here it is set where the
 * vehicle's footprint is free at every heading by the certificate the
 * planner uses itself: no obstacle pixel (class cost < 0, or a class outside
 * the table) within Chebyshev distance ceil(hypot(mx, mz)) + 1 of the pixel,
 * mx = min_dist_x / 2, mz = min_dist_z / 2 (truncated), pixels outside the
 * frame counting as free.  Every footprint sample lies within that distance
 * of the centre whatever the heading (rotation keeps it within hypot(mx,
 * mz), truncation adds less than one pixel).
 */
#include <hip/hip_runtime.h>

#include "fr_engine.h"

__global__ void __launch_bounds__(256) k_frame_distance_to_goal(float *xyz, int W, int H, int gx, int gz)
{
    const long n = (long)W * H;
    for (long pos = (long)blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += (long)gridDim.x * blockDim.x) {
        const int z = (int)(pos / W), x = (int)(pos - (long)z * W);
        const int dx = gx - x, dz = gz - z;
        xyz[3 * pos + 1] = sqrtf((float)(dx * dx + dz * dz));
    }
}

void fr_launch_frame_distance_to_goal(float *d_xyz, int W, int H, int gx, int gz, hipStream_t st)
{
    const long n = (long)W * H;
    const long b = (n + 255) / 256;
    hipLaunchKernelGGL(k_frame_distance_to_goal, dim3((unsigned)(b < 8192 ? b : 8192)), dim3(256), 0, st, d_xyz, W, H,
                       gx, gz);
}

/* Row pass of the safe zone: per pixel, whether an obstacle pixel lies in its
 * row within horizontal distance r (one thread per row, a sliding count). */
__global__ void __launch_bounds__(256) k_frame_safe_rows(const float *xyz, const float *costs, int n_classes, int W,
                                                          int H, int r, uint8_t *near)
{
    for (int z = blockIdx.x * blockDim.x + threadIdx.x; z < H; z += gridDim.x * blockDim.x) {
        const float *row = xyz + 3 * (long)z * W;
        auto obst = [&](int x) {
            const int c = (int)row[3 * x];
            return (c < 0 || c >= n_classes || costs[c] < 0.0f) ? 1 : 0;
        };
        int cnt = 0;
        for (int x = 0; x < r && x < W; ++x)
            cnt += obst(x);
        for (int x = 0; x < W; ++x) {
            if (x + r < W)
                cnt += obst(x + r);
            near[(long)z * W + x] = cnt > 0;
            if (x - r >= 0)
                cnt -= obst(x - r);
        }
    }
}

/* Column pass: the bit where no row within vertical distance r has an
 * obstacle within r (one thread per column). */
__global__ void __launch_bounds__(256) k_frame_safe_cols(float *xyz, int W, int H, int r, const uint8_t *near)
{
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < W; x += gridDim.x * blockDim.x) {
        int cnt = 0;
        for (int z = 0; z < r && z < H; ++z)
            cnt += near[(long)z * W + x];
        for (int z = 0; z < H; ++z) {
            if (z + r < H)
                cnt += near[(long)(z + r) * W + x];
            float *t = xyz + 3 * ((long)z * W + x) + 2;
            const int tv = (int)*t;
            *t = (float)(cnt == 0 ? (tv | 0x100) : (tv & ~0x100));
            if (z - r >= 0)
                cnt -= near[(long)(z - r) * W + x];
        }
    }
}

void fr_launch_frame_safe_zone(float *d_xyz, const float *d_costs, int n_classes, int W, int H, int mx, int mz,
                               uint8_t *d_near, hipStream_t st)
{
    const int r = (int)ceil(sqrt((double)mx * mx + (double)mz * mz)) + 1;
    hipLaunchKernelGGL(k_frame_safe_rows, dim3((H + 255) / 256), dim3(256), 0, st, d_xyz, d_costs, n_classes, W, H,
                       r, d_near);
    hipLaunchKernelGGL(k_frame_safe_cols, dim3((W + 255) / 256), dim3(256), 0, st, d_xyz, W, H, r, d_near);
}
