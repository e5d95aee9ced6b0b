"""Installs variantdbscan_amd (with its built lib/libSharedVDBSCAN.so and the
vdbscan command in bin/) and the reference's package name ``vdbscan`` on top
of it. Build the library first: python -c "import __graft_entry__ as g; g.build()".
The reference's NTHREADS build knob (OpenMP search threads) has no meaning
for the GPU build and is ignored."""
from setuptools import setup

setup(
    name="variantdbscan_amd",
    version="1.0.0",
    description="VariantDBSCAN hot path on AMD Instinct MI355X",
    packages=["variantdbscan_amd", "vdbscan"],
    package_dir={"vdbscan": "variantdbscan_amd/compat/vdbscan"},
    package_data={"variantdbscan_amd": ["lib/libSharedVDBSCAN.so", "bin/vdbscan",
                                        "compat/variantdbscan.py"]},
    py_modules=[],
    install_requires=["numpy"],
    zip_safe=False,
)
