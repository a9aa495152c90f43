// Work-queue support shared by single-system and batched programs.
//
// A work-queue launch claims queue positions with an atomic add on a counter
// that must be zero when the launch starts.  The counters are cleared by this
// kernel rather than by a memset so that a sequence of work-queue launches can
// be captured into a graph and replayed.
#ifndef CLODE_QUEUE_H
#define CLODE_QUEUE_H

extern "C" __global__ void __launch_bounds__(64) clode_queue_reset(unsigned int *counters, unsigned int n)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        counters[i] = 0u;
}

#endif  // CLODE_QUEUE_H
