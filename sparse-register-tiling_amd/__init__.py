"""Sparse-register-tiling SpMM for AMD Instinct MI355X (gfx950).

C = A_sparse . B_dense with A packed into M_r-row panels whose columns are
grouped by nonzero pattern, executed by generated CDNA4 HIP kernels.

The package directory name contains hyphens; import it with
``importlib.import_module("sparse-register-tiling_amd")``.

Public surface (each mirrors a reference interface):
  MatMulSpecialized, TileConfig, SpMM_SOP, get_executor_id   (matmul)
  Acc, SOPModule, SOPTile, make_sop_module, execute_tile,
  sort_and_encode_patterns, pattern_code                      (sop)
  csr.executor, csr.spmm                                      (csr)
  precision.PanelPlan, precision.spmm  (BF16 / FP8 on MFMA 16-row panels)
  serving.Server, serving.Request      (resident weights, continuous batching)
  cnn.CnnFp32, cnn.CnnPanels, cnn.GroupedPlans, cnn.GroupedPanels
                                       (many layers in one grouped launch)
  SOPOrig, save_mapping, load_mapping, named_mapping, Mapping  (mapping)
  matrix_io.read_smtx / read_mtx / read_filelist               (file formats)
  pattern_select.select_patterns / specialized_mapping         (set-cover pattern sets)
"""
from . import mapping, matrix_io, pattern_select, workloads
from .mapping import (Mapping, SOPOrig, gen_all_combos, load_mapping, mapping_id, named_mapping,
                      read_mapping_file, save_mapping)

# Host-only helpers above import without the native library.  Everything
# below needs libsrt.so and torch; the import fails loudly if it is absent.
from . import _lib
from ._lib import LibraryNotBuilt, SrtError
from .native import NativeMapping, executors, resolve_mapping
from .matmul import MatMulSpecialized, SpMM_SOP, TileConfig, get_executor_id
from .sop import (Acc, SOPModule, SOPTile, execute_tile, make_sop_module, panel_pattern_codes,
                  pattern_code, sort_and_encode_patterns)
from . import autotune, cnn, csr, precision, serving

__version__ = "0.1.0"
