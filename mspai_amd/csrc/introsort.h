/*
 * Order-exact emulation of libstdc++'s std::sort for the (idx, rho) pairs of
 * the augmentation step.
 *
 * The reference sorts the candidate scores with
 *   std::sort(rhos, rhos + len, RHO_Comparator())      src/Spai_Sub.imp:1337-1342
 * (comparator a.rho < b.rho, src/Matrix.h:147-150) and keeps the first
 * maxnew entries (src/Spai_Sub.imp:1106-1110).  When several candidates share
 * the score at the cut, which of them are kept depends on where std::sort
 * leaves equal elements.  That is fixed by libstdc++'s algorithm: introsort
 * (median-of-three pivot moved to the front, unguarded Hoare partition,
 * depth limit 2*floor(log2 n), heap sort below it) over runs longer than 16,
 * followed by insertion sort on the first 16 and unguarded insertion on the
 * rest.  This header restates that algorithm so the GPU can reproduce the
 * reference's selection on ties; it is compiled both for the device and for
 * host-side tests that compare it with the host's std::sort.
 */
#ifndef MSPAI_INTROSORT_H
#define MSPAI_INTROSORT_H

#if defined(__HIPCC__)
#define MSPAI_HD __host__ __device__
#else
#define MSPAI_HD
#endif

namespace mspai {

struct RhoPair {
    int idx;
    double rho;
};

namespace introsort_detail {

constexpr long kThreshold = 16;

MSPAI_HD inline bool lt(const RhoPair& a, const RhoPair& b) { return a.rho < b.rho; }

MSPAI_HD inline void swap_at(RhoPair* a, RhoPair* b)
{
    RhoPair t = *a;
    *a = *b;
    *b = t;
}

MSPAI_HD inline long floor_log2(long n)
{
    long r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

MSPAI_HD inline void push_heap(RhoPair* first, long hole, long top, RhoPair value)
{
    long parent = (hole - 1) / 2;
    while (hole > top && lt(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

MSPAI_HD inline void adjust_heap(RhoPair* first, long hole, long len, RhoPair value)
{
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (lt(first[child], first[child - 1])) --child;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    push_heap(first, hole, top, value);
}

MSPAI_HD inline void heap_sort(RhoPair* first, RhoPair* last)
{
    long len = last - first;
    if (len >= 2) {
        long parent = (len - 2) / 2;
        while (true) {
            RhoPair v = first[parent];
            adjust_heap(first, parent, len, v);
            if (parent == 0) break;
            --parent;
        }
    }
    while (last - first > 1) {
        --last;
        RhoPair v = *last;
        *last = *first;
        adjust_heap(first, 0, last - first, v);
    }
}

MSPAI_HD inline void median_to_first(RhoPair* result, RhoPair* a, RhoPair* b, RhoPair* c)
{
    if (lt(*a, *b)) {
        if (lt(*b, *c))
            swap_at(result, b);
        else if (lt(*a, *c))
            swap_at(result, c);
        else
            swap_at(result, a);
    } else if (lt(*a, *c)) {
        swap_at(result, a);
    } else if (lt(*b, *c)) {
        swap_at(result, c);
    } else {
        swap_at(result, b);
    }
}

MSPAI_HD inline RhoPair* unguarded_partition(RhoPair* first, RhoPair* last, RhoPair* pivot)
{
    while (true) {
        while (lt(*first, *pivot)) ++first;
        --last;
        while (lt(*pivot, *last)) --last;
        if (!(first < last)) return first;
        swap_at(first, last);
        ++first;
    }
}

MSPAI_HD inline void unguarded_linear_insert(RhoPair* last)
{
    RhoPair v = *last;
    RhoPair* next = last - 1;
    while (lt(v, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = v;
}

MSPAI_HD inline void insertion_sort(RhoPair* first, RhoPair* last)
{
    if (first == last) return;
    for (RhoPair* i = first + 1; i != last; ++i) {
        if (lt(*i, *first)) {
            RhoPair v = *i;
            for (RhoPair* p = i; p != first; --p) *p = *(p - 1);
            *first = v;
        } else {
            unguarded_linear_insert(i);
        }
    }
}

/* The recursion on the right part of each partition is turned into an
 * explicit stack; the order in which parts are finished does not change the
 * result because the parts are disjoint. */
MSPAI_HD inline void introsort_loop(RhoPair* first, RhoPair* last, long depth)
{
    struct Frame {
        RhoPair* first;
        RhoPair* last;
        long depth;
    };
    Frame stack[64];
    int sp = 0;
    stack[sp++] = {first, last, depth};
    while (sp > 0) {
        Frame f = stack[--sp];
        RhoPair* lo = f.first;
        RhoPair* hi = f.last;
        long d = f.depth;
        while (hi - lo > kThreshold) {
            if (d == 0) {
                heap_sort(lo, hi);
                break;
            }
            --d;
            RhoPair* mid = lo + (hi - lo) / 2;
            median_to_first(lo, lo + 1, mid, hi - 1);
            RhoPair* cut = unguarded_partition(lo + 1, hi, lo);
            stack[sp++] = {cut, hi, d};
            hi = cut;
        }
    }
}

}  // namespace introsort_detail

/* Sort [first, last) exactly as libstdc++'s std::sort with a.rho < b.rho. */
MSPAI_HD inline void std_sort_rho(RhoPair* first, RhoPair* last)
{
    using namespace introsort_detail;
    if (first == last) return;
    introsort_loop(first, last, floor_log2(last - first) * 2);
    if (last - first > kThreshold) {
        insertion_sort(first, first + kThreshold);
        for (RhoPair* i = first + kThreshold; i != last; ++i) unguarded_linear_insert(i);
    } else {
        insertion_sort(first, last);
    }
}

}  // namespace mspai

#endif
