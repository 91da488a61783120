"""ctypes mirrors of include/cr/uniforms.h and include/cr/abi.h.

The uniform structs are byte-identical to the reference's
CloudRendering-Vulkan/UniformBuffers.h (sizes asserted at import).
"""
import ctypes as C

f32, u32, i32, f64, u64, u16 = C.c_float, C.c_uint32, C.c_int32, C.c_double, C.c_uint64, C.c_uint16


class PushConstants(C.Structure):  # UniformBuffers.h:6-13
    _fields_ = [("time", f64), ("seed", i32), ("frameCount", u32), ("pmRadius", f32), ("currentBuffer", u32)]


class CameraProperties(C.Structure):  # UniformBuffers.h:15-69
    _fields_ = [("position", f32 * 3), ("halfWidth", i32), ("forward", f32 * 3), ("halfHeight", i32),
                ("right", f32 * 3), ("nearPlane", f32), ("up", f32 * 3), ("pixelSizeY", f32), ("pixelSizeX", f32)]


class CloudProperties(C.Structure):  # UniformBuffers.h:71-78
    _fields_ = [("bounds", (f32 * 4) * 2), ("voxelCount", u32 * 4), ("maxExtinction", f32),
                ("baseScaling", f32), ("densityScaling", f32)]


class Parameters(C.Structure):  # UniformBuffers.h:80-112
    _fields_ = [("maxRayBounces", u32), ("lightIntensity", f32), ("phaseG", f32), ("phaseOnePlusG2", f32),
                ("phaseOneMinusG2", f32), ("phaseOneOver2G", f32), ("isotropic", u32)]


class ShadowVolumeProperties(C.Structure):  # UniformBuffers.h:114-167
    _fields_ = [("bounds", (f32 * 4) * 2), ("lightDirection", f32 * 4), ("right", f32 * 4), ("up", f32 * 4),
                ("basisChange", f32 * 16), ("voxelAxisCount", u32), ("voxelSize", f32)]


class PhotonBeam(C.Structure):  # UniformBuffers.h:169-177
    _fields_ = [("startPos", f32 * 3), ("mortonCode", u32), ("endPos", f32 * 3), ("radius", f32),
                ("dataIdx", u32), ("_padding_beam", u32 * 3)]


class TreeNode(C.Structure):  # UniformBuffers.h:179-188
    _fields_ = [("parent", u32), ("left", u32), ("right", u32), ("isLeaf", u32), ("bounds", (f32 * 4) * 2),
                ("processed", u32), ("_padding_node", u32 * 3)]


class PhotonBeamData(C.Structure):  # UniformBuffers.h:190-194
    _fields_ = [("power", f32 * 4), ("trDistances", f32 * 16)]


class Photon(C.Structure):  # UniformBuffers.h:202-210
    _fields_ = [("position", f32 * 4), ("power", f32 * 4), ("phi", f32), ("theta", f32), ("_padding", f32 * 2)]


class PhotonMapProperties(C.Structure):  # UniformBuffers.h:212-249
    _fields_ = [("lightDirection", f32 * 4), ("bounds", (f32 * 4) * 2), ("voxelCount", i32 * 3), ("voxelSize", f32),
                ("photonSize", u32), ("stepSize", f32), ("absorption", f32)]


class PpmParams(C.Structure):
    _fields_ = [("numPaths", u32), ("sunProbability", f32), ("ambientStep", f32), ("maxCollisions", u32)]


class SortElement(C.Structure):  # UniformBuffers.h:196-200
    _fields_ = [("idx", u32), ("mortonCode", u32)]


class LBVHPushConstants(C.Structure):  # RenderTechniquePPB.h:41-47
    _fields_ = [("baseShift", u32), ("currentBuffer", u32), ("workGroupCount", u32)]


class GridHeader(C.Structure):
    _fields_ = [("sizeX", u32), ("sizeY", u32), ("sizeZ", u32), ("_pad", u32),
                ("voxelSizeX", f64), ("voxelSizeY", f64), ("voxelSizeZ", f64)]


class FbmParams(C.Structure):
    _fields_ = [("size", u32 * 3), ("voxelSize", f32), ("seed", u32), ("octaves", u32), ("lacunarity", f32),
                ("gain", f32), ("baseFrequency", f32), ("shape", u32), ("thresholdLo", f32), ("thresholdHi", f32)]


class RenderExt(C.Structure):
    _fields_ = [("albedo", f32), ("bounceCap", u32), ("skyMode", u32), ("skyColor", f32 * 4), ("neeMode", u32),
                ("trackingMode", u32), ("svDims", u32 * 3), ("majorantBrick", u32), ("_reserved", u32 * 3)]


class PtStats(C.Structure):
    _fields_ = [("paths", u64), ("vertices", u64), ("lookups", u64), ("svLookups", u64), ("brickVisits", u64),
                ("lastKernelMs", f64), ("totalKernelMs", f64), ("launches", u64)]


class BeamTraceParams(C.Structure):
    _fields_ = [("numPaths", u32), ("seed", i32), ("radius", f32), ("albedo", f32), ("lightDirection", f32 * 4),
                ("lightIntensity", f32), ("maxInteractions", u32), ("_reserved", u32 * 2)]


class GatherExt(C.Structure):
    _fields_ = [("refCompat", u32), ("jitter", u32), ("albedo", f32), ("_reserved", u32)]


class GatherStats(C.Structure):
    _fields_ = [("rays", u64), ("nodesVisited", u64), ("beamsTested", u64), ("lastKernelMs", f64)]


CR_BATCH_MAX_JOBS = 64
CR_PT_STATIC_MAX_FRAMES = 4096
CR_PPM_PHOTONS_PER_CELL = 32
CR_PPM_ALPHA = 0.8  # RenderTechniquePPM.h:63 (a float32 constant in the C header)
CR_PT_STATIC_SCRATCH_BYTES = 256 + 8 * CR_PT_STATIC_MAX_FRAMES
CR_PPB_ALPHA = 0.8  # RenderTechniquePPB.h:115 (a float32 constant in the C header)
CR_PPB_BEAMS_PER_PASS = 56  # RenderTechniquePPB.h:118-120


class ProgressiveState(C.Structure):
    _fields_ = [("frameCount", u32), ("framesRendered", u32), ("restarts", u32), ("randState", u32)]


class BatchTuning(C.Structure):
    _fields_ = [("blockThreads", u32), ("blocksPerCU", u32), ("tileWidth", u32), ("_reserved", u32)]


class BatchTuneReport(C.Structure):
    _fields_ = [("defaultMs", f32), ("bestMs", f32), ("candidates", u32), ("jobBrick", u32 * CR_BATCH_MAX_JOBS)]


SIZES = {
    PushConstants: 24, CameraProperties: 68, CloudProperties: 60, Parameters: 28, ShadowVolumeProperties: 152,
    PhotonBeam: 48, TreeNode: 64, PhotonBeamData: 80, SortElement: 8, LBVHPushConstants: 12,
}
for _t, _n in SIZES.items():
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)

# status codes and enums (abi.h)
CR_OK, CR_ERR_INVALID_ARGUMENT, CR_ERR_OUT_OF_MEMORY, CR_ERR_DEVICE = 0, 1, 2, 3
CR_ERR_CAPACITY, CR_ERR_IO, CR_ERR_STATE, CR_ERR_NOT_SUPPORTED = 4, 5, 6, 7
CR_UNBOUNDED = 0xFFFFFFFF
CR_DTYPE_F32, CR_DTYPE_F16 = 0, 1
CR_LAYOUT_X_FASTEST, CR_LAYOUT_Z_FASTEST = 0, 1
CR_GRID_DENSE, CR_GRID_BRICK4 = 0, 1
CR_SHAPE_CUMULUS, CR_SHAPE_ANVIL = 0, 1
CR_SKY_GRADIENT, CR_SKY_CONSTANT = 0, 1
CR_NEE_SHADOW_VOLUME, CR_NEE_RATIO_TRACKING, CR_NEE_NONE = 0, 1, 2
CR_TRACK_GLOBAL_MAJORANT, CR_TRACK_BRICK_DDA = 0, 1
