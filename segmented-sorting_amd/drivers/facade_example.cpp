// The cub-style C++ facade (include/segsort.hpp) on the example printed in cub's
// documentation (lib/cub/device/device_segmented_radix_sort.cuh:104-129, 302-327,
// 506-526, 685-705): all eight overloads (SortPairs, SortPairsDescending,
// SortKeys, SortKeysDescending, each with separate in/out pointers and with
// DoubleBuffers) are called the way cub's are, for every algorithm, and their
// output compared with the documented answer. Exit status 0 when all match.
#include <cstdio>
#include <vector>

#include "segsort.hpp"

namespace {

int failures = 0;

void expect(const char* what, const std::vector<int>& got, const std::vector<int>& want) {
    if (got != want) {
        ++failures;
        printf("%s: mismatch\n", what);
    }
}

template <typename T>
std::vector<int> fetch(const T* d, int n) {
    std::vector<int> h((size_t)n);
    if (hipMemcpy(h.data(), d, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) ++failures;
    return h;
}

}  // namespace

int main() {
    const int num_items = 7, num_segments = 3;
    const std::vector<int> offsets = {0, 3, 3, 7}, keys = {8, 6, 7, 5, 3, 0, 9}, vals = {0, 1, 2, 3, 4, 5, 6};
    int *d_offsets, *d_keys_in, *d_keys_out, *d_values_in, *d_values_out;
    if (hipMalloc((void**)&d_offsets, 4 * sizeof(int)) || hipMalloc((void**)&d_keys_in, 7 * sizeof(int)) ||
        hipMalloc((void**)&d_keys_out, 7 * sizeof(int)) || hipMalloc((void**)&d_values_in, 7 * sizeof(int)) ||
        hipMalloc((void**)&d_values_out, 7 * sizeof(int)))
        return 2;
    (void)hipMemcpy(d_offsets, offsets.data(), 4 * sizeof(int), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_keys_in, keys.data(), 7 * sizeof(int), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_values_in, vals.data(), 7 * sizeof(int), hipMemcpyHostToDevice);

    for (int algo = SEGSORT_ALGO_BATCHED; algo <= SEGSORT_ALGO_FIXPASS; ++algo) {
        void* d_temp = nullptr;
        size_t temp_bytes = 0;
        using segsort::DeviceSegmentedRadixSort;
        DeviceSegmentedRadixSort::SortPairs(d_temp, temp_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out,
                                            num_items, num_segments, d_offsets, d_offsets + 1, 0, 32, 0, false, algo);
        if (hipMalloc(&d_temp, temp_bytes + 4096) != hipSuccess) return 2;
        temp_bytes += 4096;  // room for every overload below
        DeviceSegmentedRadixSort::SortPairs(d_temp, temp_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out,
                                            num_items, num_segments, d_offsets, d_offsets + 1, 0, 32, 0, true, algo);
        expect("SortPairs keys", fetch(d_keys_out, 7), {6, 7, 8, 0, 3, 5, 9});
        expect("SortPairs values", fetch(d_values_out, 7), {1, 2, 0, 5, 4, 3, 6});
        DeviceSegmentedRadixSort::SortPairsDescending(d_temp, temp_bytes, d_keys_in, d_keys_out, d_values_in,
                                                      d_values_out, num_items, num_segments, d_offsets,
                                                      d_offsets + 1, 0, 32, 0, true, algo);
        expect("SortPairsDescending keys", fetch(d_keys_out, 7), {8, 7, 6, 9, 5, 3, 0});
        expect("SortPairsDescending values", fetch(d_values_out, 7), {0, 2, 1, 6, 3, 4, 5});
        DeviceSegmentedRadixSort::SortKeys(d_temp, temp_bytes, d_keys_in, d_keys_out, num_items, num_segments,
                                           d_offsets, d_offsets + 1, 0, 32, 0, true, algo);
        expect("SortKeys", fetch(d_keys_out, 7), {6, 7, 8, 0, 3, 5, 9});
        DeviceSegmentedRadixSort::SortKeysDescending(d_temp, temp_bytes, d_keys_in, d_keys_out, num_items,
                                                     num_segments, d_offsets, d_offsets + 1, 0, 32, 0, true, algo);
        expect("SortKeysDescending", fetch(d_keys_out, 7), {8, 7, 6, 9, 5, 3, 0});
        // DoubleBuffer forms: the result ends in the alternate buffer and the selector flips
        int *d_kalt, *d_valt;
        if (hipMalloc((void**)&d_kalt, 7 * sizeof(int)) || hipMalloc((void**)&d_valt, 7 * sizeof(int))) return 2;
        auto reset = [&] {
            (void)hipMemcpy(d_keys_out, keys.data(), 7 * sizeof(int), hipMemcpyHostToDevice);
            (void)hipMemcpy(d_values_out, vals.data(), 7 * sizeof(int), hipMemcpyHostToDevice);
        };
        auto flipped = [&](int sel) {
            if (sel != 1) ++failures;
        };
        {
            reset();
            segsort::DoubleBuffer<int> dk(d_keys_out, d_kalt), dv(d_values_out, d_valt);
            DeviceSegmentedRadixSort::SortPairs(d_temp, temp_bytes, dk, dv, num_items, num_segments, d_offsets,
                                                d_offsets + 1, 0, 32, 0, true, algo);
            flipped(dk.selector);
            flipped(dv.selector);
            expect("DoubleBuffer SortPairs keys", fetch(dk.Current(), 7), {6, 7, 8, 0, 3, 5, 9});
            expect("DoubleBuffer SortPairs values", fetch(dv.Current(), 7), {1, 2, 0, 5, 4, 3, 6});
        }
        {
            reset();
            segsort::DoubleBuffer<int> dk(d_keys_out, d_kalt), dv(d_values_out, d_valt);
            DeviceSegmentedRadixSort::SortPairsDescending(d_temp, temp_bytes, dk, dv, num_items, num_segments,
                                                          d_offsets, d_offsets + 1, 0, 32, 0, true, algo);
            flipped(dk.selector);
            flipped(dv.selector);
            expect("DoubleBuffer SortPairsDescending keys", fetch(dk.Current(), 7), {8, 7, 6, 9, 5, 3, 0});
            expect("DoubleBuffer SortPairsDescending values", fetch(dv.Current(), 7), {0, 2, 1, 6, 3, 4, 5});
        }
        {
            reset();
            segsort::DoubleBuffer<int> dk(d_keys_out, d_kalt);
            DeviceSegmentedRadixSort::SortKeys(d_temp, temp_bytes, dk, num_items, num_segments, d_offsets,
                                               d_offsets + 1, 0, 32, 0, true, algo);
            flipped(dk.selector);
            expect("DoubleBuffer SortKeys", fetch(dk.Current(), 7), {6, 7, 8, 0, 3, 5, 9});
        }
        {
            reset();
            segsort::DoubleBuffer<int> dk(d_keys_out, d_kalt);
            DeviceSegmentedRadixSort::SortKeysDescending(d_temp, temp_bytes, dk, num_items, num_segments, d_offsets,
                                                         d_offsets + 1, 0, 32, 0, true, algo);
            flipped(dk.selector);
            expect("DoubleBuffer SortKeysDescending", fetch(dk.Current(), 7), {8, 7, 6, 9, 5, 3, 0});
        }
        (void)hipFree(d_kalt);
        (void)hipFree(d_valt);
        (void)hipFree(d_temp);
    }
    printf(failures ? "facade FAILED (%d)\n" : "facade ok\n", failures);
    return failures ? 1 : 0;
}
