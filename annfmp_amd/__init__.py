"""MI355X-native propagation-assisted k-d-tree nearest-neighbour fields.

Public API (mirrors diku-dk/annfmp's openmp backend):
  ANNField                 -- front with algorithm="hip"   (annfmp/base.py)
  ANNFieldPropKDTreeHIP    -- backend class               (annfmp/openmp/base.py)
  KDTree                   -- tree wrapper                (annfmp/openmp/kdtree/base.py)
The compute path is libannf_hip.so (C ABI: include/annf_hip.h).
"""
from .field import ANNField
from .base import ANNFieldPropKDTreeHIP, field_on_device, field_mm_on_device
from .kdtree import KDTree

__all__ = ["ANNField", "ANNFieldPropKDTreeHIP", "KDTree", "field_on_device", "field_mm_on_device"]
