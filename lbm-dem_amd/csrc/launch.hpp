// Kernel launches for one box or for a batch of boxes.
//
// Every kernel is written once, as a __device__ body. kernel_one passes the body's arguments as
// kernel arguments (one box); kernel_batch reads the arguments of box blockIdx.y from a device
// table, so one launch advances every box of a batch (grid.x: the largest per-box grid; a body
// whose work depends on its grid takes the per-box grid as an argument and lets surplus blocks
// return). launch<> either launches kernel_one or, while a Recorder is installed on the calling
// thread, appends the launch (kernel, grid, block, argument bytes) to it; a batch records the same
// step on every box and replays the records with kernel_batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cxxabi.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <utility>
#include <vector>

namespace lbdem {

// the arguments of one launch: an aggregate of trivially copyable members (copied as bytes)
template <size_t I, typename T>
struct PackLeaf {
    T v;
};
template <typename Seq, typename... P>
struct PackImpl;
template <size_t... I, typename... P>
struct PackImpl<std::index_sequence<I...>, P...> : PackLeaf<I, P>... {};
template <typename... P>
using Pack = PackImpl<std::index_sequence_for<P...>, P...>;

// The table is read through the constant address space (it does not change while a kernel runs):
// pointers loaded from there are known to address global memory, as pointers in kernel arguments
// are, so the bodies keep their global (not flat) memory instructions and atomic combining.
#define LBDEM_CONST __attribute__((address_space(4)))
template <size_t I, typename T>
__device__ inline T leaf_value(const LBDEM_CONST PackLeaf<I, T>& l) {
    if constexpr (std::is_class<T>::value) return *(const T*)&l.v;  // trivially copyable struct
    else return l.v;
}
template <auto Body, typename... P, size_t... I>
__device__ inline void apply_pack(const LBDEM_CONST PackImpl<std::index_sequence<I...>, P...>& t) {
    Body(leaf_value<I, P>(t)...);
}

template <auto Body, int MAXT, int MINW, typename... P>
__global__ __launch_bounds__(MAXT, MINW) void kernel_one(P... args) {
    Body(args...);
}

template <auto Body, int MAXT, int MINW, typename... P>
__global__ __launch_bounds__(MAXT, MINW) void kernel_batch(const Pack<P...>* tab) {
    const LBDEM_CONST Pack<P...>* ct = (const LBDEM_CONST Pack<P...>*)tab;
    apply_pack<Body, P...>(ct[blockIdx.y]);
}

// one recorded launch of one box
struct RecordedLaunch {
    const void* key;  // identity of the kernel (body and bounds)
    const char* name; // the launcher's signature, naming the body (for timing tables)
    void (*batch)(dim3 grid, dim3 block, size_t shm, hipStream_t s, const void* table);
    dim3 grid, block;
    size_t shm;
    unsigned elastic;                 // grid-stride body: blocks over all boxes of a batch (0: exact grid)
    std::vector<unsigned char> args;  // the body's argument tuple, as bytes
};
struct Recorder {
    std::vector<RecordedLaunch> ops;
};
Recorder*& current_recorder();  // thread-local; null: launches execute

template <auto Body, int MAXT, int MINW, typename Sig = decltype(Body)>
struct Launcher;
template <auto Body, int MAXT, int MINW, typename... P>
struct Launcher<Body, MAXT, MINW, void (*)(P...)> {
    using Tuple = Pack<P...>;
    static_assert(std::is_trivially_copyable<Tuple>::value, "kernel arguments are copied as bytes");
    static inline const char tag = 0;
    static const char* signature() {  // "lbdem::Launcher<&lbdem::k_fused<float, 0, true, true, 8>, 256, 4, ...>"
        static const std::string s = [] {
            int st = 0;
            char* d = abi::__cxa_demangle(typeid(Launcher).name(), nullptr, nullptr, &st);
            std::string r = (st == 0 && d) ? d : typeid(Launcher).name();
            std::free(d);
            return r;
        }();
        return s.c_str();
    }
    static void batch(dim3 g, dim3 b, size_t shm, hipStream_t s, const void* table) {
        hipLaunchKernelGGL((kernel_batch<Body, MAXT, MINW, P...>), g, b, shm, s, static_cast<const Tuple*>(table));
    }
    template <typename... A>
    static void run(unsigned elastic, dim3 g, dim3 b, size_t shm, hipStream_t s, A&&... a) {
        if (Recorder* r = current_recorder()) {
            const Tuple t{{static_cast<P>(std::forward<A>(a))}...};
            RecordedLaunch op{&tag, signature(), &batch, g, b, shm, elastic, std::vector<unsigned char>(sizeof(Tuple))};
            std::memcpy(op.args.data(), &t, sizeof(Tuple));
            r->ops.push_back(std::move(op));
            return;
        }
        hipLaunchKernelGGL((kernel_one<Body, MAXT, MINW, P...>), g, b, shm, s, static_cast<P>(std::forward<A>(a))...);
    }
};

// launch<&body, max threads per block, min waves per SIMD>(grid, block, stream, args...)
template <auto Body, int MAXT = 1024, int MINW = 1, typename... A>
inline void launch(dim3 grid, dim3 block, hipStream_t s, A&&... a) {
    Launcher<Body, MAXT, MINW>::run(0u, grid, block, 0, s, std::forward<A>(a)...);
}
// the same for a grid-stride body, whose grid a batch may shrink: it spreads BLOCKS blocks over
// its boxes (kernels that usually return at once take few; kernels with work in most steps more)
template <auto Body, int MAXT = 1024, int MINW = 1, unsigned BLOCKS = 2048, typename... A>
inline void launch_gs(dim3 grid, dim3 block, hipStream_t s, A&&... a) {
    Launcher<Body, MAXT, MINW>::run(BLOCKS, grid, block, 0, s, std::forward<A>(a)...);
}

}  // namespace lbdem
