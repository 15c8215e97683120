// fixcub: the reference's src/fixcub.cu driver over this library. The three
// steps of the reference's timed region (DeviceReduce::Max, adjustment<Plus>,
// DeviceRadixSort::SortPairs, adjustment<Minus>; fixcub.cu:140-166) are one
// fix-pass call here, with the segment id packed into a 64-bit composite so it
// cannot overflow. Temp is allocated once, outside the timed region.
#include "driver_io.hpp"
#include "segsort.hpp"

int main(int argc, char** argv) {
    drv::Options o = drv::parse(argc, argv, SEGSORT_ALGO_FIXPASS);
    std::vector<int> seg;
    std::vector<unsigned> keys;
    if (!drv::read_input(seg, keys)) {
        fprintf(stderr, "fixcub: malformed input\n");
        return EXIT_FAILURE;
    }
    const int S = (int)seg.size() - 1, N = (int)keys.size();
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
    drv::check(hipMemcpy(d_seg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice), "H2D");
    drv::check(hipMemcpy(d_val, vals.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
    hipEvent_t start, stop;
    drv::check(hipEventCreate(&start), "hipEventCreate");
    drv::check(hipEventCreate(&stop), "hipEventCreate");
    size_t temp_bytes = 0;
    drv::check(segsort::DeviceSegmentedRadixSort::SortPairs(nullptr, temp_bytes, d_vec, d_vec_out, d_val, d_val_out,
                                                            N, S, d_seg, d_seg + 1, 0, 32, 0, false, o.algo),
               "temp query");
    void* d_temp = nullptr;
    drv::check(hipMalloc(&d_temp, temp_bytes), "hipMalloc temp");
    float total_ms = 0;
    for (int ex = -o.warmup; ex < o.executions; ++ex) {  // ex < 0: untimed warm-up calls
        drv::check(hipMemcpy(d_vec, keys.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipEventRecord(start, 0), "hipEventRecord");
        drv::check(segsort::DeviceSegmentedRadixSort::SortPairs(d_temp, temp_bytes, d_vec, d_vec_out, d_val,
                                                                d_val_out, N, S, d_seg, d_seg + 1, 0, 32, 0, false,
                                                                o.algo),
                   "fix-pass SortPairs");
        drv::check(hipEventRecord(stop, 0), "hipEventRecord");
        drv::check(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0;
        drv::check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (ex >= 0) total_ms += ms;
    }
    std::vector<unsigned> out((size_t)N);
    drv::check(hipMemcpy(out.data(), d_vec_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    drv::check(hipMemcpy(vals.data(), d_val_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.time) printf("%f\n", total_ms / o.executions);
    else {
        drv::print_keys(out.data(), out.size());
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
