/*
 * dcg_program_plugin.hpp -- build a vertex-program plugin for the engine.
 *
 * A plugin is a HIP source that defines one or more programs on
 * dcg::api::AppBase (include/dcg_program.hpp) and exports their program table:
 *
 *     #include "dcg_program_plugin.hpp"
 *     struct MyProgram : dcg::api::AppBase<uint32_t, uint32_t, uint32_t> { ... };
 *     DCG_DEFINE_PROGRAM(my_program_table, MyProgram)
 *
 * built with the engine's flags into a shared library:
 *
 *     hipcc --offload-arch=gfx950 -O3 -fPIC -shared -std=c++17 \
 *           -Xclang -target-feature -Xclang -packed-fp32-ops \
 *           -I include my_program.hip -o libmy_program.so
 *
 * The engine's repair kernels (dycgraph_amd/csrc/vertex_kernels.hpp) are
 * instantiated for the programs inside the plugin; the exported function returns
 * a dcg_program table that dcg_set_program installs before a graph load.  This is
 * the run-time counterpart of instantiating sepgraph::engine::Engine<..., TApp>
 * with a new app struct (framework.cuh:780-783).
 */
#ifndef DCG_PROGRAM_PLUGIN_HPP
#define DCG_PROGRAM_PLUGIN_HPP

#include "dcg_program.hpp"
#include "../dycgraph_amd/csrc/vertex_kernels.hpp"

// extern "C" const dcg_program *NAME(void): the table of the program set
// ProgramSet<__VA_ARGS__> (program i = query_apps value i).
#define DCG_DEFINE_PROGRAM(NAME, ...)                                                            \
    extern "C" const dcg_program *NAME(void) {                                                   \
        static const dcg_program table = ::dcg::ProgramTable<::dcg::ProgramSet<__VA_ARGS__>>::make(#NAME); \
        return &table;                                                                           \
    }

#endif
