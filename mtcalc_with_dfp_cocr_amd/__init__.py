"""MI355X-native COCR solver for the reference's 3-D magnetotelluric edge-element systems.

Public surface:
  model      regular-mesh models, assembly (native), reference file formats
  cocr       host mirror of PCOCR::PCOCR_2x2_Folded and the handle-based batched solver
  build      in-tree build of the HIP library and the drop-in COCR_FP executable
"""
__all__ = ["model", "cocr", "build"]
