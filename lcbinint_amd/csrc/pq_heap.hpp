// Binary max-heap of (key, payload) pairs whose pops come out in exactly the
// order of a libstdc++ std::priority_queue<T> (std::less on the key) fed the
// same pushes: push is std::push_heap's hole walk up from the back, pop is
// std::pop_heap's hole walk down from the root along the larger child (the
// left one on a tie between children), with the last element then walked up
// from the hole, as libstdc++'s __adjust_heap / __push_heap do.  Ties between
// equal keys are therefore broken as the reference's refinement queue breaks
// them (LocalRefinementQueueItem, finite_source_magnifier.cpp:969-976, used at
// :3202-3283).
//
// The heap lives in caller-provided arrays (LDS or device memory on the GPU,
// plain arrays on the host); `Store` is the store used for every element
// write, so that device code can write uniform values through its wave store.
#pragma once

#ifndef LCBG_HD
#if defined(__HIPCC__)
#define LCBG_HD __host__ __device__
#else
#define LCBG_HD
#endif
#endif

namespace lcbg {

struct PlainStore {
    template <class T>
    LCBG_HD void operator()(T* p, T v) const { *p = v; }
};

template <class Store = PlainStore>
struct PqHeap {
    double* key;
    int* val;
    int n;
    Store put;

    // std::__push_heap(first, hole, top = 0, value)
    LCBG_HD void sift_up(int hole, double k, int v)
    {
        int parent = (hole - 1) / 2;
        while (hole > 0 && key[parent] < k) {
            put(&key[hole], key[parent]);
            put(&val[hole], val[parent]);
            hole = parent;
            parent = (hole - 1) / 2;
        }
        put(&key[hole], k);
        put(&val[hole], v);
    }

    LCBG_HD void push(double k, int v)
    {
        const int hole = n;
        n = n + 1;
        sift_up(hole, k, v);
    }

    // Removes the largest element (n > 0) and returns its payload and key.
    LCBG_HD int pop(double* top_key)
    {
        const double tk = key[0];
        const int tv = val[0];
        const int len = n - 1;  // std::pop_heap then pop_back
        if (len > 0) {
            const double lk = key[len];
            const int lv = val[len];
            // std::__adjust_heap(first, 0, len, last)
            int hole = 0;
            int second = 0;
            while (second < (len - 1) / 2) {
                second = 2 * (second + 1);
                if (key[second] < key[second - 1]) --second;
                put(&key[hole], key[second]);
                put(&val[hole], val[second]);
                hole = second;
            }
            if ((len & 1) == 0 && second == (len - 2) / 2) {
                second = 2 * (second + 1);
                put(&key[hole], key[second - 1]);
                put(&val[hole], val[second - 1]);
                hole = second - 1;
            }
            sift_up(hole, lk, lv);
        }
        n = len;
        *top_key = tk;
        return tv;
    }
};

}  // namespace lcbg
