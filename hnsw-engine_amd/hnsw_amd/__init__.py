"""hnsw_amd: the hnsw-engine search path on AMD Instinct MI355X (gfx950).

Drop-in for the reference's search API -- ``HNSWIndex.search`` / ``batch_search(filter=)``,
``QuantizedIndex.batch_search(rerank=)``, ``exact_search`` / ``exact_distances`` -- over graphs
in the reference's ``save_native`` layout. Searches run hand-written HIP kernels through the C
ABI of ``libhnsw_mi355x.so`` (``include/hnsw_mi355x.h``); graph construction is sequential
insertion on the host, identical to the reference's.
"""

from .exact import exact_distances, exact_search
from .graph_io import GraphState
from .index import ENGINES, SELECTIONS, HNSWIndex, IndexParams
from .metrics import METRICS
from .quantization import QuantizedIndex, ScalarQuantizer
from .recall import recall_at_k

__all__ = [
    "ENGINES", "METRICS", "SELECTIONS", "GraphState", "HNSWIndex", "IndexParams",
    "QuantizedIndex", "ScalarQuantizer", "exact_distances", "exact_search", "recall_at_k",
]
