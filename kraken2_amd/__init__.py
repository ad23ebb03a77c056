"""MI355X-native Kraken 2 classify core.

The hot path (minimizer scan, compact-hash probe, hit-count/LCA resolution)
runs as HIP kernels for gfx950 inside libk2g.so; this package is the thin
Python binding used by tests and the benchmark.  The C ABI is
include/k2g.h; the command-line front-ends are bin/k2g_classify and
bin/k2g_merge (multi-database merge of kraken outputs).
"""
from ._lib import (  # noqa: F401
    BIN_DIR,
    CODE_AMBIG,
    CODE_FRAME,
    CODE_MATE,
    EXPORTED_SYMBOLS,
    LIB_PATH,
    Batch,
    Builder,
    BuildParams,
    ClassifyOutput,
    Db,
    K2GError,
    Merger,
    MergeOutput,
    Pipeline,
    Server,
    expand_spaced_seed_mask,
    format_hitlist,
    load_library,
    pack_fragments,
    tune_pipeline,
    write_report,
)
