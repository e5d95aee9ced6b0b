"""``vdbscan.variantdbscan`` of the reference (python_package/vdbscan/variantdbscan.py)."""
from variantdbscan_amd.variantdbscan import VariantDBSCAN  # noqa: F401

__all__ = ["VariantDBSCAN"]
