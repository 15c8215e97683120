// nseg: the reference's per-segment multi-stream baselines (src/ncub.cu,
// src/newthrust.cu) over this library. stdin -> device -> timed
// segsort_multistream_sort (keys only, one device-wide sort per segment) ->
// stdout. ncub.cu:113-145 deals the segments over NUM_STREAMS = 4 streams;
// newthrust.cu:91-167 over up to 32, round robin or with -DSCHEDULE to the
// least-loaded stream. Here: --streams K (default 4) and --schedule; with
// --values the sort carries the input index as value and prints it too.
#include "driver_io.hpp"
#include "segsort.h"

int main(int argc, char** argv) {
    drv::Options o = drv::parse(argc, argv, SEGSORT_ALGO_BATCHED);
    int streams = 4, schedule = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--streams") && i + 1 < argc) streams = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--schedule")) schedule = 1;
    }
    std::vector<int> seg;
    std::vector<unsigned> keys;
    if (!drv::read_input(seg, keys)) {
        fprintf(stderr, "nseg: malformed input\n");
        return EXIT_FAILURE;
    }
    const int S = (int)seg.size() - 1, N = (int)keys.size();
    std::vector<unsigned> vals((size_t)N);
    for (int i = 0; i < N; ++i) vals[(size_t)i] = (unsigned)i;
    const int vb = o.values ? 4 : 0;
    unsigned *d_vec, *d_vec_out, *d_val = nullptr, *d_val_out = nullptr;
    const size_t nb = sizeof(unsigned) * (size_t)(N > 0 ? N : 1);
    drv::check(hipMalloc((void**)&d_vec, nb), "hipMalloc");
    drv::check(hipMalloc((void**)&d_vec_out, nb), "hipMalloc");
    if (o.values) {
        drv::check(hipMalloc((void**)&d_val, nb), "hipMalloc");
        drv::check(hipMalloc((void**)&d_val_out, nb), "hipMalloc");
    }
    hipEvent_t start, stop;
    drv::check(hipEventCreate(&start), "hipEventCreate");
    drv::check(hipEventCreate(&stop), "hipEventCreate");
    size_t temp_bytes = 0;
    if (segsort_multistream_sort(nullptr, &temp_bytes, nullptr, nullptr, nullptr, nullptr, N, S, seg.data(), 0, 32,
                                 SEGSORT_KEY_U32, vb, 0, streams, schedule, nullptr) != 0) {
        fprintf(stderr, "nseg: %s\n", segsort_last_error());
        return EXIT_FAILURE;
    }
    void* d_temp = nullptr;
    drv::check(hipMalloc(&d_temp, temp_bytes), "hipMalloc temp");
    float total_ms = 0;
    for (int ex = -o.warmup; ex < o.executions; ++ex) {  // ex < 0: untimed warm-up calls
        drv::check(hipMemcpy(d_vec, keys.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        drv::check(hipMemcpy(d_vec_out, keys.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        if (o.values) {
            drv::check(hipMemcpy(d_val, vals.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
            drv::check(hipMemcpy(d_val_out, vals.data(), sizeof(unsigned) * (size_t)N, hipMemcpyHostToDevice), "H2D");
        }
        drv::check(hipEventRecord(start, 0), "hipEventRecord");
        if (segsort_multistream_sort(d_temp, &temp_bytes, d_vec, d_vec_out, d_val, d_val_out, N, S, seg.data(), 0,
                                     32, SEGSORT_KEY_U32, vb, 0, streams, schedule, nullptr) != 0) {
            fprintf(stderr, "nseg: %s\n", segsort_last_error());
            return EXIT_FAILURE;
        }
        drv::check(hipEventRecord(stop, 0), "hipEventRecord");
        drv::check(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0;
        drv::check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        if (ex >= 0) total_ms += ms;
    }
    drv::check(hipMemcpy(keys.data(), d_vec_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.values)
        drv::check(hipMemcpy(vals.data(), d_val_out, sizeof(unsigned) * (size_t)N, hipMemcpyDeviceToHost), "D2H");
    if (o.time) printf("%f\n", total_ms / o.executions);
    else {
        drv::print_keys(keys.data(), keys.size());
        if (o.values) drv::print_values(vals.data(), vals.size());
    }
    (void)hipFree(d_temp);
    (void)hipFree(d_val);
    (void)hipFree(d_val_out);
    (void)hipFree(d_vec);
    (void)hipFree(d_vec_out);
    return 0;
}
