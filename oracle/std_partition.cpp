// TEST INFRASTRUCTURE ONLY: libstdc++'s std::partition, called the way the
// reference's SAH build calls it (Src/Core/bvh.cu:159-165: raw element
// pointers, predicate "bucket <= best split"), so tests can check the oracle's
// triangle order against the standard library's own algorithm instead of
// against a second restatement of it.  Raw pointers are random-access
// iterators; libstdc++ dispatches them to its bidirectional (two-pointer swap)
// partition, exactly as it does for the reference's Triangle*.
#include <algorithm>
#include <cstdint>

extern "C" {

// Partition ids[0..n) in place by bucket[ids[k]] <= best; returns the size of
// the true zone.
int std_partition_ids(int32_t* ids, int n, const int32_t* bucket, int best) {
    int32_t* mid = std::partition(ids, ids + n, [bucket, best](int32_t id) { return bucket[id] <= best; });
    return (int)(mid - ids);
}

// libstdc++'s version macro, so a test can name the library it checked against.
int std_partition_glibcxx(void) {
#ifdef __GLIBCXX__
    return (int)__GLIBCXX__;
#else
    return 0;
#endif
}
}
