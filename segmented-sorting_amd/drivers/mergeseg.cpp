// mergeseg: the reference's src/mergeseg.cu driver over this library.
// mgpu::segmented_sort(d_vec, d_index, N, d_seg, S, less, context)
// (mergeseg.cu:118) becomes segsort_mgpu_segmented_sort: begin offsets only,
// keys and values sorted in place. The index values are filled with 0..N-1
// (the reference passes an uninitialised buffer, mergeseg.cu:105).
#include "driver_io.hpp"
#include "segsort.h"

int main(int argc, char** argv) {
    drv::Options o = drv::parse(argc, argv, SEGSORT_ALGO_BATCHED);
    std::vector<int> seg;
    std::vector<unsigned> keys;
    if (!drv::read_input(seg, keys)) {
        fprintf(stderr, "mergeseg: malformed input\n");
        return EXIT_FAILURE;
    }
    const int S = (int)seg.size() - 1, N = (int)keys.size();  // the last offset is dropped (mergeseg.cu:79-89)
    std::vector<unsigned> idx((size_t)N);
    for (int i = 0; i < N; ++i) idx[(size_t)i] = (unsigned)i;
    unsigned *d_vec, *d_idx;
    int* d_seg;
    const size_t nb = sizeof(unsigned) * (size_t)(N > 0 ? N : 1);
    drv::check(hipMalloc((void**)&d_seg, sizeof(int) * (size_t)(S > 0 ? S : 1)), "hipMalloc");
    drv::check(hipMalloc((void**)&d_vec, nb), "hipMalloc");
    drv::check(hipMalloc((void**)&d_idx, nb), "hipMalloc");
    hipEvent_t start, stop;
    drv::check(hipEventCreate(&start), "hipEventCreate");
    drv::check(hipEventCreate(&stop), "hipEventCreate");
    float total_ms = 0;
    for (int ex = -o.warmup; ex < o.executions; ++ex) {  // ex < 0: untimed warm-up calls
        if (S > 0) drv::check(hipMemcpy(d_seg, seg.data(), sizeof(int) * (size_t)S, hipMemcpyHostToDevice), "H2D");
        drv::check(hipMemcpy(d_vec, keys.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipMemcpy(d_idx, idx.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipEventRecord(start, 0), "hipEventRecord");
        const int rc = segsort_mgpu_segmented_sort(d_vec, d_idx, N, d_seg, S, SEGSORT_KEY_U32, 4, nullptr);
        if (rc != 0) {
            fprintf(stderr, "segmented_sort failed: %s\n", segsort_last_error());
            return EXIT_FAILURE;
        }
        drv::check(hipEventRecord(stop, 0), "hipEventRecord");
        drv::check(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0;
        drv::check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (ex >= 0) total_ms += ms;
    }
    std::vector<unsigned> out((size_t)N);
    drv::check(hipMemcpy(out.data(), d_vec, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    drv::check(hipMemcpy(idx.data(), d_idx, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.time) printf("%f\n", total_ms / o.executions);
    else {
        drv::print_keys(out.data(), out.size());
        if (o.values) drv::print_values(idx.data(), idx.size());
    }
    (void)hipFree(d_seg);
    (void)hipFree(d_vec);
    (void)hipFree(d_idx);
    return 0;
}
