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
