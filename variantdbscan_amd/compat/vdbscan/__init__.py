"""The reference's Python package name (python_package/setup.py: package
``vdbscan``, module ``vdbscan.variantdbscan``), served by variantdbscan_amd:
``from vdbscan.variantdbscan import VariantDBSCAN`` and
``from vdbscan import VariantDBSCAN`` both give the drop-in call."""
from variantdbscan_amd.variantdbscan import VariantDBSCAN  # noqa: F401

__all__ = ["VariantDBSCAN"]
