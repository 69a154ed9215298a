"""MI355X-native incremental streaming-graph engine (SSSP/BFS/CC/PageRank delta repair).

The compute path is libdcg.so (hand-written gfx950 HIP kernels behind a C ABI,
include/dcg.h); this package is the host-side mirror of the reference engine's
interface over that ABI.
"""
from .engine import (APP_BFS, APP_CC, APP_PR, APP_SSSP, CACHE_HOTNESS, CACHE_LEARNED, CACHE_NONE, INF,
                     PLACE_DEVICE, PLACE_HOST, W_REFERENCE, W_SEEDED, DcgError, Engine, EngineConfig,
                     gen_rmat, gen_stream, gen_grid, gen_lognormal_sizes, calibrate, calibrate_lines, mlp_eval, lib,
                     LIB_PATH, load_program_plugin, radix_sort, read_graph, write_graph, read_updates,
                     write_updates)

__all__ = ["Engine", "EngineConfig", "DcgError", "gen_rmat", "gen_stream", "gen_grid", "gen_lognormal_sizes", "calibrate", "calibrate_lines", "mlp_eval", "lib", "LIB_PATH", "load_program_plugin", "radix_sort",
           "read_graph", "write_graph", "read_updates", "write_updates",
           "APP_SSSP", "APP_BFS", "APP_CC", "APP_PR", "W_REFERENCE", "W_SEEDED", "PLACE_DEVICE",
           "PLACE_HOST", "CACHE_NONE", "CACHE_HOTNESS", "CACHE_LEARNED", "INF"]
