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

/* Element moves are written field by field: a whole-struct copy becomes a
 * private-memory temporary in device code. */
MSPAI_HD inline void move_to(RhoPair* dst, const RhoPair* src)
{
    dst->idx = src->idx;
    dst->rho = src->rho;
}

MSPAI_HD inline void swap_at(RhoPair* a, RhoPair* b)
{
    int ti = a->idx;
    double tr = a->rho;
    move_to(a, b);
    b->idx = ti;
    b->rho = tr;
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

MSPAI_HD inline void push_heap(RhoPair* first, long hole, long top, int vi, double vr)
{
    long parent = (hole - 1) / 2;
    while (hole > top && first[parent].rho < vr) {
        move_to(&first[hole], &first[parent]);
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole].idx = vi;
    first[hole].rho = vr;
}

MSPAI_HD inline void adjust_heap(RhoPair* first, long hole, long len, int vi, double vr)
{
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (lt(first[child], first[child - 1])) --child;
        move_to(&first[hole], &first[child]);
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        move_to(&first[hole], &first[child - 1]);
        hole = child - 1;
    }
    push_heap(first, hole, top, vi, vr);
}

MSPAI_HD inline void heap_sort(RhoPair* first, RhoPair* last)
{
    long len = last - first;
    if (len >= 2) {
        long parent = (len - 2) / 2;
        while (true) {
            adjust_heap(first, parent, len, first[parent].idx, first[parent].rho);
            if (parent == 0) break;
            --parent;
        }
    }
    while (last - first > 1) {
        --last;
        int vi = last->idx;
        double vr = last->rho;
        move_to(last, first);
        adjust_heap(first, 0, last - first, vi, vr);
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
    int vi = last->idx;
    double vr = last->rho;
    RhoPair* next = last - 1;
    while (vr < next->rho) {
        move_to(last, next);
        last = next;
        --next;
    }
    last->idx = vi;
    last->rho = vr;
}

MSPAI_HD inline void insertion_sort(RhoPair* first, RhoPair* last)
{
    if (first == last) return;
    for (RhoPair* i = first + 1; i != last; ++i) {
        if (lt(*i, *first)) {
            int vi = i->idx;
            double vr = i->rho;
            for (RhoPair* p = i; p != first; --p) move_to(p, p - 1);
            first->idx = vi;
            first->rho = vr;
        } else {
            unguarded_linear_insert(i);
        }
    }
}

/* The recursion on the right part of each partition is turned into an
 * explicit stack of (first, last, depth) index triples held in `stack`
 * (kStackFrames frames, caller-provided so device code keeps it out of
 * private memory).  Frames on the stack have strictly decreasing depth, so at
 * most 2*floor(log2 n) + 1 are live.  The order in which parts are finished
 * does not change the result because the parts are disjoint. */
constexpr int kStackFrames = 64;

MSPAI_HD inline void introsort_loop(RhoPair* base, long n, long depth, int* stack)
{
    int sp = 0;
    stack[0] = 0;
    stack[1] = (int)n;
    stack[2] = (int)depth;
    sp = 1;
    while (sp > 0) {
        --sp;
        long lo = stack[3 * sp], hi = stack[3 * sp + 1], d = stack[3 * sp + 2];
        while (hi - lo > kThreshold) {
            if (d == 0) {
                heap_sort(base + lo, base + hi);
                break;
            }
            --d;
            long mid = lo + (hi - lo) / 2;
            median_to_first(base + lo, base + lo + 1, base + mid, base + hi - 1);
            long cut = unguarded_partition(base + lo + 1, base + hi, base + lo) - base;
            stack[3 * sp] = (int)cut;
            stack[3 * sp + 1] = (int)hi;
            stack[3 * sp + 2] = (int)d;
            ++sp;
            hi = cut;
        }
    }
}

}  // namespace introsort_detail

/* Sort [first, last) exactly as libstdc++'s std::sort with a.rho < b.rho.
 * `stack` holds 3 * introsort_detail::kStackFrames ints of scratch. */
MSPAI_HD inline void std_sort_rho(RhoPair* first, RhoPair* last, int* stack)
{
    using namespace introsort_detail;
    if (first == last) return;
    introsort_loop(first, last - first, floor_log2(last - first) * 2, stack);
    if (last - first > kThreshold) {
        insertion_sort(first, first + kThreshold);
        for (RhoPair* i = first + kThreshold; i != last; ++i) unguarded_linear_insert(i);
    } else {
        insertion_sort(first, last);
    }
}

}  // namespace mspai

#endif
