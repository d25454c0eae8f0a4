"""braincell_amd: an MI355X-native staggered cable solver for braincell-style
multi-compartment cells (hand-written HIP kernels behind a C ABI)."""

__version__ = "0.1.0"
