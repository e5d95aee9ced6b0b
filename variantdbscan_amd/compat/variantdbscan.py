"""Top-level ``variantdbscan`` module, as the reference's examples import it
(src/vtest.py, src/variant_example.py: ``from variantdbscan import
VariantDBSCAN`` next to src/variantdbscan.py). Put this directory on the
path to run them unchanged."""
from variantdbscan_amd.variantdbscan import VariantDBSCAN  # noqa: F401

__all__ = ["VariantDBSCAN"]
