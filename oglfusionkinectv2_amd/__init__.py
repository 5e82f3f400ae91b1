"""MI355X (gfx950) implementation of the oglFusionKinectV2 per-frame
KinectFusion loop: HIP kernels behind the C ABI of include/kfusion.h, with a
host mirror of the reference's gFusion class."""
from . import _lib
from .gfusion import GFusion, GFusionConfig, kinect_config, translate, KINECT_V2_CAM
from .mcubes import MCubes, MCubeConfig

__all__ = ["GFusion", "GFusionConfig", "kinect_config", "translate", "KINECT_V2_CAM", "MCubes", "MCubeConfig",
           "_lib"]
