"""VariantDBSCAN hot path on AMD Instinct MI355X.

Grid-indexed eps-neighbour search and DBSCAN cluster expansion over many
(eps, minpts) variants, in hand-written HIP kernels for gfx950 behind the
reference's C ABI (``libVDBSCAN`` in ``lib/libSharedVDBSCAN.so``) and its
Python call (``VariantDBSCAN``).
"""
from .variantdbscan import VariantDBSCAN  # noqa: F401
from . import _lib  # noqa: F401
from ._lib import (cluster, cluster_device, cluster_multi, cluster_multi_device,  # noqa: F401
                   neighbor_counts, neighbor_table, table_digest,
                   rank_order, PATH_AUTO, PATH_FUSED, PATH_TABLE, PATH_BATCH, PREC_FP64,
                   PREC_FP32_RECHECK, BORDER_REFERENCE, BORDER_STANDARD, LIB_PATH)

__all__ = ["VariantDBSCAN", "cluster", "cluster_device", "cluster_multi", "neighbor_counts",
           "neighbor_table", "rank_order", "table_digest"]
