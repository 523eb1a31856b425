// Kernel bodies and segmented launches.
//
// Every kernel is written as a body, `struct K { LB, MB; static __device__ void
// run(KGrid, args...) }`, that takes its block index and grid size from a KGrid
// instead of blockIdx.x / gridDim.x.  `kl<K>(grid, block, lds, stream, args...)`
// launches it for one cluster (kone: KGrid is the hardware grid).
//
// While a SegRecorder is installed on the calling thread, kl records the launch
// (kernel, grid, block, LDS bytes, argument bytes) instead of issuing it.  An
// ensemble then merges the recorded launches of many clusters into one kseg launch
// per kernel: its blocks are the clusters' grids laid end to end (each starting at a
// multiple of 8 blocks, so a block keeps the XCD its cluster-local index implies);
// a block finds its cluster in a table indexed by its block index over 8 and reads
// that cluster's grid and arguments from device memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <type_traits>
#include <utility>
#include <vector>

struct KGrid { uint32_t bid, nb, by, nby; };

#define CMCG_AS1 __attribute__((address_space(1)))

// An argument as kept in device memory for a merged launch.  Pointers, also those
// inside structs, are typed as global-address-space pointers there and turned back
// into ordinary pointers when read, so the compiler keeps global (not flat) memory
// instructions for them; the bits are unchanged (global and flat addresses agree).
template <typename T, typename = void>
struct SegArg {
    using type = T;
    static __device__ __forceinline__ T get(const T &x) { return x; }
};
template <typename T>
struct SegArg<T *, void> {
    using type = CMCG_AS1 T *;
    static __device__ __forceinline__ T *get(CMCG_AS1 T *p) { return (T *)p; }
};
template <typename T>
struct SegArg<T, std::enable_if_t<std::is_class<T>::value && sizeof(T) % 8 == 0 && alignof(T) >= 8>> {
    struct alignas(alignof(T)) type { uint64_t w[sizeof(T) / 8]; };
    static __device__ __forceinline__ T get(const type &x)
    {
        T out;
#pragma unroll
        for (unsigned i = 0; i < sizeof(T) / 8; i++) {
            char *g = (char *)(CMCG_AS1 char *)x.w[i];
            memcpy((char *)&out + 8 * i, &g, 8);
        }
        return out;
    }
};

// A plain tuple with the same layout on host and device.
template <typename... P> struct Tup;
template <> struct Tup<> {};
template <typename H, typename... T> struct Tup<H, T...> { H h; Tup<T...> t; };

template <size_t I> struct TupAt {
    template <typename H, typename... T>
    static __host__ __device__ __forceinline__ const auto &get(const Tup<H, T...> &x) { return TupAt<I - 1>::get(x.t); }
};
template <> struct TupAt<0> {
    template <typename H, typename... T>
    static __host__ __device__ __forceinline__ const H &get(const Tup<H, T...> &x) { return x.h; }
};

inline Tup<> tup_make() { return {}; }
template <typename H, typename... T>
inline Tup<H, T...> tup_make(H h, T... t) { return Tup<H, T...>{h, tup_make(t...)}; }

template <typename Body, typename... P, size_t... I>
__device__ __forceinline__ void seg_apply(KGrid g, const Tup<typename SegArg<P>::type...> &a, std::index_sequence<I...>)
{
    Body::run(g, SegArg<P>::get(TupAt<I>::get(a))...);
}

// ent[c] = (first block, blocks) of cluster c; chunk[b / 8] = the cluster that owns
// blocks 8 (b / 8) .. 8 (b / 8) + 7 (cluster grids start at multiples of 8)
template <typename Body, typename... P>
__device__ __forceinline__ void seg_run(const uint2 *__restrict__ ent, const uint16_t *__restrict__ chunk,
                                        const Tup<typename SegArg<P>::type...> *__restrict__ a)
{
    const uint32_t b = blockIdx.x;
    const int c = chunk[b >> 3];
    const uint2 e = ent[c];
    if (b - e.x >= e.y) return;             // alignment padding after the cluster's grid
    seg_apply<Body, P...>(KGrid{b - e.x, e.y, blockIdx.y, gridDim.y}, a[c], std::index_sequence_for<P...>{});
}

template <typename Body, int LB, typename... P>
__global__ __launch_bounds__(LB) void kone(P... p) { Body::run(KGrid{blockIdx.x, gridDim.x, blockIdx.y, gridDim.y}, p...); }
template <typename Body, int LB, int MB, typename... P>
__global__ __launch_bounds__(LB, MB) void kone_mb(P... p)
{
    Body::run(KGrid{blockIdx.x, gridDim.x, blockIdx.y, gridDim.y}, p...);
}
template <typename Body, int LB, typename... P>
__global__ __launch_bounds__(LB) void kseg(const uint2 *__restrict__ ent, const uint16_t *__restrict__ chunk,
                                           const Tup<typename SegArg<P>::type...> *__restrict__ a)
{
    seg_run<Body, P...>(ent, chunk, a);
}
template <typename Body, int LB, int MB, typename... P>
__global__ __launch_bounds__(LB, MB) void kseg_mb(const uint2 *__restrict__ ent, const uint16_t *__restrict__ chunk,
                                                  const Tup<typename SegArg<P>::type...> *__restrict__ a)
{
    seg_run<Body, P...>(ent, chunk, a);
}

// ---------------------------------------------------------------- host side
// Launches a merged kernel: `blocks` blocks in x (the clusters' aligned grids), gy in y.
using SegLaunchFn = void (*)(uint32_t blocks, uint32_t gy, dim3 block, size_t lds, hipStream_t st, const uint2 *ent,
                             const uint16_t *chunk, const void *args);

struct SegRec {
    SegLaunchFn fn;                  // also identifies the kernel instantiation
    uint32_t gx, gy;
    dim3 block;
    size_t lds;
    std::vector<unsigned char> args; // Tup<P...> bytes
};

struct SegRecorder { std::vector<SegRec> recs; bool bad = false; };   // bad: a launch that cannot be merged

// set while a cluster's step is recorded instead of launched (calling thread only)
extern thread_local SegRecorder *g_seg_rec;

template <typename Body, typename... P>
void seg_launch(uint32_t blocks, uint32_t gy, dim3 block, size_t lds, hipStream_t st, const uint2 *ent,
                const uint16_t *chunk, const void *args)
{
    static_assert(sizeof(Tup<typename SegArg<P>::type...>) == sizeof(Tup<P...>), "argument layout");
    using A = const Tup<typename SegArg<P>::type...> *;
    if constexpr (Body::MB > 0)
        hipLaunchKernelGGL((kseg_mb<Body, Body::LB, Body::MB, P...>), dim3(blocks, gy), block, lds, st, ent, chunk, (A)args);
    else
        hipLaunchKernelGGL((kseg<Body, Body::LB, P...>), dim3(blocks, gy), block, lds, st, ent, chunk, (A)args);
}

template <typename T> inline T seg_conv(T x) { return x; }

template <typename Body, typename... P, typename... A>
inline void kl_impl(void (*)(KGrid, P...), dim3 g, dim3 b, size_t lds, hipStream_t st, A &&...a)
{
    if (SegRecorder *r = g_seg_rec) {
        if (g.x == 0 || g.y == 0 || g.z == 0) return;
        if (g.z != 1) { r->bad = true; return; }
        const Tup<P...> t = tup_make<P...>(seg_conv<P>(std::forward<A>(a))...);
        SegRec rec{&seg_launch<Body, P...>, g.x, g.y, b, lds, std::vector<unsigned char>(sizeof(t))};
        memcpy(rec.args.data(), &t, sizeof(t));
        r->recs.push_back(std::move(rec));
        return;
    }
    if constexpr (Body::MB > 0)
        hipLaunchKernelGGL((kone_mb<Body, Body::LB, Body::MB, P...>), g, b, lds, st, seg_conv<P>(std::forward<A>(a))...);
    else
        hipLaunchKernelGGL((kone<Body, Body::LB, P...>), g, b, lds, st, seg_conv<P>(std::forward<A>(a))...);
}

// Launches kernel body K on `grid` (or records it, see above).
template <typename Body, typename... A>
inline void kl(dim3 g, dim3 b, size_t lds, hipStream_t st, A &&...a)
{
    kl_impl<Body>(decltype(&Body::run){}, g, b, lds, st, std::forward<A>(a)...);
}
