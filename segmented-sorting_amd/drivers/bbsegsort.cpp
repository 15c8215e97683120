// bbsegsort: the reference's src/bb_segsort/main.cu driver over this library.
// int keys with double values equal to the keys (main.cu:59-72), sorted in
// place by bb_segsort(key_d, val_d, N, seg_d, S) (main.cu:102), here
// segsort_bb_segsort, which is stable. The S+1 offsets are read into an
// array of S+1 entries (the reference's vector of S overflows, main.cu:61-62).
#include "driver_io.hpp"
#include "segsort.h"

int main(int argc, char** argv) {
    drv::Options o = drv::parse(argc, argv, SEGSORT_ALGO_BATCHED);
    std::vector<int> seg;
    std::vector<unsigned> ukeys;
    if (!drv::read_input(seg, ukeys)) {
        fprintf(stderr, "bbsegsort: malformed input\n");
        return EXIT_FAILURE;
    }
    const int S = (int)seg.size() - 1, N = (int)ukeys.size();
    std::vector<int> h_vec((size_t)N);
    std::vector<double> h_val((size_t)N);
    for (int i = 0; i < N; ++i) {
        h_vec[(size_t)i] = (int)ukeys[(size_t)i];
        h_val[(size_t)i] = h_vec[(size_t)i];
    }
    int *key_d, *seg_d;
    double* val_d;
    drv::check(hipMalloc((void**)&key_d, sizeof(int) * (size_t)(N > 0 ? N : 1)), "hipMalloc");
    drv::check(hipMalloc((void**)&val_d, sizeof(double) * (size_t)(N > 0 ? N : 1)), "hipMalloc");
    drv::check(hipMalloc((void**)&seg_d, sizeof(int) * (size_t)(S > 0 ? S : 1)), "hipMalloc");
    if (S > 0) drv::check(hipMemcpy(seg_d, seg.data(), sizeof(int) * (size_t)S, hipMemcpyHostToDevice), "H2D");
    hipEvent_t start, stop;
    drv::check(hipEventCreate(&start), "hipEventCreate");
    drv::check(hipEventCreate(&stop), "hipEventCreate");
    float total_ms = 0;
    std::vector<int> out = h_vec;
    for (int ex = -o.warmup; ex < o.executions; ++ex) {  // ex < 0: untimed warm-up calls
        drv::check(hipMemcpy(key_d, h_vec.data(), sizeof(int) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipMemcpy(val_d, h_val.data(), sizeof(double) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipEventRecord(start, 0), "hipEventRecord");
        if (!segsort_bb_segsort(key_d, val_d, N, seg_d, S, SEGSORT_KEY_I32, 8, nullptr)) {
            fprintf(stderr, "bb_segsort failed: %s\n", segsort_last_error());
            return EXIT_FAILURE;
        }
        drv::check(hipEventRecord(stop, 0), "hipEventRecord");
        drv::check(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0;
        drv::check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (ex >= 0) total_ms += ms;
    }
    drv::check(hipMemcpy(out.data(), key_d, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    std::vector<double> vout((size_t)N);
    drv::check(hipMemcpy(vout.data(), val_d, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.time) printf("%f\n", total_ms / o.executions);
    else {
        drv::print_keys(out.data(), out.size());
        if (o.values) drv::print_values(vout.data(), vout.size());  // the values are the keys, as doubles
    }
    (void)hipFree(key_d);
    (void)hipFree(val_d);
    (void)hipFree(seg_d);
    return 0;
}
