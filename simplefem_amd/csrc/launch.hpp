// launch.hpp -- kernel launches that can be recorded into a hipGraph.
//
// Every kernel of the library is launched through sfa::launch.  While a
// GraphRecorder is active on the calling thread, a launch becomes a kernel
// node of the recorder's graph that depends on the node recorded before it
// (the order one stream would impose) instead of running.  The graph is built
// node by node, not by stream capture, so other threads may use the runtime
// (allocate, copy, free, run other streams) while it is being built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace sfa {

struct GraphRecorder;
// The calling thread's active recorder, or null (defined in kernels.hip).
GraphRecorder*& current_recorder();

struct GraphRecorder {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    bool has_last = false;
    std::vector<std::shared_ptr<void>> keep;  // argument storage of the recorded nodes
    hipError_t error = hipSuccess;            // first failure while recording

    GraphRecorder()
    {
        error = hipGraphCreate(&graph, 0);
        current_recorder() = this;
    }
    ~GraphRecorder()
    {
        if (current_recorder() == this) current_recorder() = nullptr;
        if (graph) (void)hipGraphDestroy(graph);
    }
    GraphRecorder(const GraphRecorder&) = delete;
    GraphRecorder& operator=(const GraphRecorder&) = delete;

    // Stops recording and instantiates the graph.
    hipError_t instantiate(hipGraphExec_t* exec)
    {
        current_recorder() = nullptr;
        if (error != hipSuccess) return error;
        return hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    }
};

namespace detail {

template <typename Tuple, std::size_t... I>
void fill_pointers(Tuple& t, void** ptrs, std::index_sequence<I...>)
{
    ((ptrs[I] = static_cast<void*>(&std::get<I>(t))), ...);
}

}  // namespace detail

template <typename... P, typename... A>
void launch(void (*kernel)(P...), dim3 grid, dim3 block, hipStream_t stream, A&&... args)
{
    static_assert(sizeof...(P) == sizeof...(A), "argument count");
    GraphRecorder* rec = current_recorder();
    if (!rec) {
        kernel<<<grid, block, 0, stream>>>(static_cast<P>(std::forward<A>(args))...);
        return;
    }
    if (rec->error != hipSuccess) return;
    auto values = std::make_shared<std::tuple<P...>>(static_cast<P>(std::forward<A>(args))...);
    void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
    detail::fill_pointers(*values, ptrs, std::index_sequence_for<P...>{});
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = ptrs;
    p.extra = nullptr;
    hipGraphNode_t node;
    rec->error = hipGraphAddKernelNode(&node, rec->graph, rec->has_last ? &rec->last : nullptr,
                                       rec->has_last ? 1 : 0, &p);
    if (rec->error != hipSuccess) return;
    rec->last = node;
    rec->has_last = true;
    rec->keep.push_back(values);
}

}  // namespace sfa
