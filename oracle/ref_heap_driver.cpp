// Thin extern "C" driver around the reference's own heap (src/heap.hpp), compiled from the
// reference source where it lies by oracle/build.py into oracle/_ref/. Used only by tests to
// pin the oracle's heap restatement against the real thing, ties included.
#include "heap.hpp"

extern "C" {
void ref_init_result(int64_t size, float* dis, int64_t* ids) {
    harmony::init_result(harmony::METRIC_L2, (size_t)size, dis, ids);
}
void ref_heap_replace_top(int64_t k, float* v, int64_t* ids, float val, int64_t id) {
    harmony::heap_replace_top<harmony::METRIC_L2>((size_t)k, v, ids, val, id);
}
void ref_sort_result(int64_t size, float* dis, int64_t* ids) {
    harmony::sort_result(harmony::METRIC_L2, (size_t)size, dis, ids);
}
}
