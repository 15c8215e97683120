// radixseg: the reference's src/radixseg.cu driver over this library.
// stdin -> device -> timed SortPairs<uint, uint> (values = index) -> stdout.
// cub::DeviceSegmentedRadixSort::SortPairs (radixseg.cu:116-124) becomes
// segsort::DeviceSegmentedRadixSort::SortPairs with the same arguments.
#include "driver_io.hpp"
#include "segsort.hpp"

#include <algorithm>

int main(int argc, char** argv) {
    drv::Options o = drv::parse(argc, argv, SEGSORT_ALGO_BATCHED);
    std::vector<int> seg;
    std::vector<unsigned> keys;
    if (!drv::read_input(seg, keys)) {
        fprintf(stderr, "radixseg: malformed input\n");
        return EXIT_FAILURE;
    }
    const int S = (int)seg.size() - 1, N = (int)keys.size();
    // --bound: the longest segment, which the host offsets give for free, as the
    // call's length bound (segsort_sort_hinted); 0 = none, cub's call
    int bound = 0;
    for (int i = 0; o.bound && i < S; ++i) bound = std::max(bound, seg[(size_t)i + 1] - seg[(size_t)i]);
    std::vector<unsigned> vals((size_t)N);
    for (int i = 0; i < N; ++i) vals[(size_t)i] = (unsigned)i;
    unsigned *d_vec, *d_vec_out, *d_val, *d_val_out;
    int* d_seg;
    const size_t nb = sizeof(unsigned) * (size_t)(N > 0 ? N : 1);
    drv::check(hipMalloc((void**)&d_seg, sizeof(int) * seg.size()), "hipMalloc");
    drv::check(hipMalloc((void**)&d_vec, nb), "hipMalloc");
    drv::check(hipMalloc((void**)&d_vec_out, nb), "hipMalloc");
    drv::check(hipMalloc((void**)&d_val, nb), "hipMalloc");
    drv::check(hipMalloc((void**)&d_val_out, nb), "hipMalloc");
    hipEvent_t start, stop;
    drv::check(hipEventCreate(&start), "hipEventCreate");
    drv::check(hipEventCreate(&stop), "hipEventCreate");
    void* d_temp = nullptr;
    size_t temp_bytes = 0;
    float total_ms = 0;
    for (int ex = -o.warmup; ex < o.executions; ++ex) {  // ex < 0: untimed warm-up calls
        // the reference uploads the inputs before every timed call (radixseg.cu:111-114); with
        // --resident they are uploaded once (the sort never modifies them), so that the call is
        // timed on inputs already resident in HBM, as bench.py times it
        if (!o.resident || ex == -o.warmup) {
            drv::check(hipMemcpy(d_seg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice), "H2D");
            drv::check(hipMemcpy(d_vec, keys.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
            drv::check(hipMemcpy(d_val, vals.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        }
        if (!d_temp) {
            drv::check(segsort::DeviceSegmentedRadixSort::SortPairs(nullptr, temp_bytes, d_vec, d_vec_out, d_val,
                                                                    d_val_out, N, S, d_seg, d_seg + 1, 0, 32, 0,
                                                                    false, o.algo, bound),
                       "temp query");
            drv::check(hipMalloc(&d_temp, temp_bytes), "hipMalloc temp");
        }
        drv::check(hipEventRecord(start, 0), "hipEventRecord");
        drv::check(segsort::DeviceSegmentedRadixSort::SortPairs(d_temp, temp_bytes, d_vec, d_vec_out, d_val,
                                                                d_val_out, N, S, d_seg, d_seg + 1, 0, 32, 0, false,
                                                                o.algo, bound),
                   "SortPairs");
        drv::check(hipEventRecord(stop, 0), "hipEventRecord");
        drv::check(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0;
        drv::check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (ex >= 0) total_ms += ms;
    }
    drv::check(hipMemcpy(keys.data(), d_vec_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    drv::check(hipMemcpy(vals.data(), d_val_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.time) printf("%f\n", total_ms / o.executions);
    else {
        drv::print_keys(keys.data(), keys.size());
        if (o.values) drv::print_values(vals.data(), vals.size());
    }
    (void)hipFree(d_temp);
    (void)hipFree(d_seg);
    (void)hipFree(d_vec);
    (void)hipFree(d_vec_out);
    (void)hipFree(d_val);
    (void)hipFree(d_val_out);
    return 0;
}
