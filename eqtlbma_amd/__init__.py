"""MI355X-native Bayes-factor engine behind eqtlbma_bf.

The product is the C ABI in include/eqbf.h, implemented by
eqtlbma_amd/libeqbf_hip.so (HIP kernels for gfx950) and the eqtlbma_bf and
eqtlbma_hm host programs.  This package holds the sources, the build recipe and a thin Python
binding over the C ABI (eqtlbma_amd.engine).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libeqbf_hip.so")
CLI_PATH = os.path.join(PKG_DIR, "eqtlbma_bf")
CLI_HM_PATH = os.path.join(PKG_DIR, "eqtlbma_hm")
CLI_AVG_PATH = os.path.join(PKG_DIR, "eqtlbma_avg_bfs")
