"""MI355X (gfx950) engine for Severo.jl's embedding-and-graph path.

normalize_cells -> scale_features -> embedding (PCA via IRLBA)
-> nearest_neighbours -> shared_nearest_neighbours / jaccard_index,
computed by hand-written HIP kernels in libsevero_hip.so behind a C ABI
(include/severo_hip.h).  This package mirrors the reference's Julia API;
``julia/SeveroHIP.jl`` is the Julia-side binding of the same library.

The directory name is not a Python identifier; load it with
``importlib`` (``load_package()`` in tests and bench.py does this) under the
module name ``severo_jl_amd``.
"""
from ._lib import ConvergenceError, SeveroError, load  # noqa: F401
from .api import (CenteredMatrix, LinearEmbedding, NamedCounts, SVD, ann, cluster, copy_probe,  # noqa: F401
                  embedding, irlba, make_unique, read_10X, read_mm,
                  find_variable_features, irlba_operator,
                  jaccard_index, nearest_neighbours, normalize_cells, pca, scale_features,
                  shared_nearest_neighbours)
from .pipeline import Pipeline  # noqa: F401
from .synth import NBTables, nb_tables, top_mean_features  # noqa: F401
