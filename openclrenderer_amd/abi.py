"""numpy views of the records declared in include/ocr.h.

The layouts are byte-compatible with the reference's host structs:
``struct vertex``/``struct triangle`` (cl2.cl:131-151, vertex.hpp:8-38,
triangle.hpp:8-15), ``struct obj_g_descriptor`` (cl2.cl:99-124,
obj_g_descriptor.hpp:9-36) and ``struct light`` (cl2.cl:77-87,
light.hpp:25-69).
"""
import numpy as np

VERTEX = np.dtype([
    ("pos", "<f4", 4),
    ("normal", "<f4", 4),
    ("vt", "<f4", 2),
    ("object_id", "<u4"),
    ("vertex_col", "<u4"),
])
assert VERTEX.itemsize == 48

TRIANGLE = np.dtype([("v", VERTEX, 3)])
assert TRIANGLE.itemsize == 144

OBJ_DESC = np.dtype([
    ("world_pos", "<f4", 4),
    ("world_rot_quat", "<f4", 4),
    ("old_world_pos_1", "<f4", 4),
    ("old_world_pos_2", "<f4", 4),
    ("old_world_rot_quat_1", "<f4", 4),
    ("old_world_rot_quat_2", "<f4", 4),
    ("scale", "<f4"),
    ("tid", "<u4"),
    ("rid", "<u4"),
    ("ssid", "<u4"),
    ("has_bump", "<u4"),
    ("specular", "<f4"),
    ("spec_mult", "<f4"),
    ("diffuse", "<f4"),
    ("buffer_offset", "<i4"),
    ("feature_flag", "<i4"),
    ("_pad", "<u4", 2),
])
assert OBJ_DESC.itemsize == 144

LIGHT = np.dtype([
    ("pos", "<f4", 4),
    ("col", "<f4", 4),
    ("shadow", "<u4"),
    ("brightness", "<f4"),
    ("radius", "<f4"),
    ("diffuse", "<f4"),
    ("godray_intensity", "<f4"),
    ("is_static", "<i4"),
    ("_pad", "<u4", 2),
])
assert LIGHT.itemsize == 64

VIEW = np.dtype([
    ("width", "<u4"),
    ("height", "<u4"),
    ("fov_const", "<f4"),
    ("use_linear_rendering", "<u4"),
    ("c_pos", "<f4", 4),
    ("c_rot", "<f4", 4),
    ("clear_col", "<f4", 4),
])
assert VIEW.itemsize == 64

FRAME_STATS = np.dtype([
    ("n_cut", "<u4"),
    ("n_frag", "<u4"),
    ("overflow", "<u4"),
    ("n_profiled", "<u4"),
    ("ms_clear", "<f4"),
    ("ms_prearrange", "<f4"),
    ("ms_depth", "<f4"),
    ("ms_id", "<f4"),
    ("ms_shade", "<f4"),
    ("ms_total", "<f4"),
])

# feature flags, cl2.cl:90-97
FEATURE_SS_REFLECTIVE = 1
FEATURE_TWO_SIDED = 2
FEATURE_OUTLINE = 4
FEATURE_IS_STATIC = 8
FEATURE_DOES_NOT_RECEIVE_DYNAMIC_SHADOWS = 16

NO_TRIANGLE = 0xFFFFFFFF
UINT_MAX = 0xFFFFFFFF

OCR_OK = 0
OCR_ERR_INVALID = -1
OCR_ERR_HIP = -2
OCR_ERR_OOM = -3
OCR_ERR_OVERFLOW = -4
OCR_ERR_UNSUPPORTED = -5
OCR_ERR_BUSY = -6
