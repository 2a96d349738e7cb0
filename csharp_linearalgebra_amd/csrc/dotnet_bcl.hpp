// Host-side restatements of the two .NET base-library routines whose exact
// behaviour fixes the norm estimator's probes (NormEstimate.cs:163, 286-305,
// 376-385): System.Random for a seeded instance and Array.Sort(keys, items).
// Plain C++ (no HIP), so the restatements can be checked on their own
// against .NET's published outputs.
#pragma once

#include <climits>
#include <cmath>
#include <cstdlib>
#include <utility>

namespace tns {

// ---------------------------------------------------------------------------
// System.Random(seed): Knuth's subtractive generator (TAOCP vol. 2, 3.6), in
// the form .NET keeps for seeded instances; Next(k) = (int)(Sample() * k).
// ---------------------------------------------------------------------------
class SeededRandom {
public:
    explicit SeededRandom(int seed)
    {
        const int mbig = INT_MAX;
        int subtraction = (seed == INT_MIN) ? INT_MAX : std::abs(seed);
        int mj = 161803398 - subtraction;
        s_[55] = mj;
        int mk = 1, ii = 0;
        for (int i = 1; i < 55; i++) {
            if ((ii += 21) >= 55) ii -= 55;
            s_[ii] = mk;
            mk = mj - mk;
            if (mk < 0) mk += mbig;
            mj = s_[ii];
        }
        for (int k = 1; k < 5; k++)
            for (int i = 1; i < 56; i++) {
                int n = i + 30;
                if (n >= 55) n -= 55;
                s_[i] -= s_[1 + n];
                if (s_[i] < 0) s_[i] += mbig;
            }
        a_ = 0;
        b_ = 21;
    }
    double sample()
    {
        if (++a_ >= 56) a_ = 1;
        if (++b_ >= 56) b_ = 1;
        int v = s_[a_] - s_[b_];
        if (v == INT_MAX) v--;
        if (v < 0) v += INT_MAX;
        s_[a_] = v;
        return v * (1.0 / INT_MAX);
    }
    int next(int max_value) { return (int)(sample() * max_value); }

private:
    int s_[56] = {0};
    int a_ = 0, b_ = 0;
};

// ---------------------------------------------------------------------------
// Array.Sort(double[] keys, int[] items): NaNs first, then introspective sort
// (median-of-three quicksort, insertion sort at <= 16, heapsort past depth
// 2*(floor(log2 n)+1)). Equal keys keep the order this procedure gives them,
// which decides which of several equal h values the estimator probes.
// ---------------------------------------------------------------------------
struct KeySort {
    double* k;
    int* v;
    void swap(int i, int j)
    {
        std::swap(k[i], k[j]);
        std::swap(v[i], v[j]);
    }
    void swap_if_greater(int i, int j)
    {
        if (k[i] > k[j]) swap(i, j);
    }
    void insertion(int lo, int n)
    {
        for (int i = 0; i < n - 1; i++) {
            double t = k[lo + i + 1];
            int tv = v[lo + i + 1];
            int j = i;
            while (j >= 0 && t < k[lo + j]) {
                k[lo + j + 1] = k[lo + j];
                v[lo + j + 1] = v[lo + j];
                j--;
            }
            k[lo + j + 1] = t;
            v[lo + j + 1] = tv;
        }
    }
    void down_heap(int lo, int i, int n)
    {
        double d = k[lo + i - 1];
        int dv = v[lo + i - 1];
        while (i <= n >> 1) {
            int child = 2 * i;
            if (child < n && k[lo + child - 1] < k[lo + child]) child++;
            if (!(d < k[lo + child - 1])) break;
            k[lo + i - 1] = k[lo + child - 1];
            v[lo + i - 1] = v[lo + child - 1];
            i = child;
        }
        k[lo + i - 1] = d;
        v[lo + i - 1] = dv;
    }
    void heap(int lo, int n)
    {
        for (int i = n >> 1; i >= 1; i--) down_heap(lo, i, n);
        for (int i = n; i > 1; i--) {
            swap(lo, lo + i - 1);
            down_heap(lo, 1, i - 1);
        }
    }
    int partition(int lo, int n)
    {
        int hi = n - 1, mid = hi >> 1;
        swap_if_greater(lo, lo + mid);
        swap_if_greater(lo, lo + hi);
        swap_if_greater(lo + mid, lo + hi);
        double pivot = k[lo + mid];
        swap(lo + mid, lo + hi - 1);
        int left = 0, right = hi - 1;
        while (left < right) {
            while (k[lo + ++left] < pivot) {
            }
            while (pivot < k[lo + --right]) {
            }
            if (left >= right) break;
            swap(lo + left, lo + right);
        }
        if (left != hi - 1) swap(lo + left, lo + hi - 1);
        return left;
    }
    void intro(int lo, int n, int depth)
    {
        int size = n;
        while (size > 1) {
            if (size <= 16) {
                if (size == 2) {
                    swap_if_greater(lo, lo + 1);
                    return;
                }
                if (size == 3) {
                    swap_if_greater(lo, lo + 1);
                    swap_if_greater(lo, lo + 2);
                    swap_if_greater(lo + 1, lo + 2);
                    return;
                }
                insertion(lo, size);
                return;
            }
            if (depth == 0) {
                heap(lo, size);
                return;
            }
            depth--;
            int p = partition(lo, size);
            intro(lo + p + 1, size - (p + 1), depth);
            size = p;
        }
    }
};

inline void sort_keys_items(int n, double* keys, int* items)
{
    if (n <= 1) return;
    int left = 0;
    for (int i = 0; i < n; i++)
        if (std::isnan(keys[i])) {
            std::swap(keys[left], keys[i]);
            std::swap(items[left], items[i]);
            left++;
        }
    int m = n - left;
    if (m <= 1) return;
    int log2n = 31 - __builtin_clz((unsigned)m);
    KeySort s{keys + left, items + left};
    s.intro(0, m, 2 * (log2n + 1));
}

}  // namespace tns
