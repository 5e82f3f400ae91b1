"""Host-side mirror of the reference's `class MCubes` (include/mcubes.h:46-170,
src/mcubes.cpp) over the surface export of the C ABI (kf_mesh_*).

The frame loop's "save stl" button (kinect.cpp:487-504) reads:

    mc = MCubes()
    cfg = MCubeConfig(); cfg.gridSize = gconfig.volumeSize
    mc.setConfig(cfg); mc.setVolumeTexture(gfusion); mc.init()
    mc.setIsolevel(0); mc.generateMarchingCubes(); mc.exportMesh()

The volume is the session's TSDF (no copy); all work runs in the HIP library.
"""
import ctypes as C

import numpy as np

from . import _lib as K


class MCubeConfig:
    """mCubeConfig (include/mcubes.h:25-43) with the reference defaults."""

    def __init__(self):
        self.gridSize = (128, 128, 128)
        self.numVoxels = 128 ** 3
        self.voxelSize = (1000.0 / 128,) * 3
        self.isoValue = 1650.0
        self.maxVerts = 128 ** 3


class MCubes:
    def __init__(self):
        self._cfg = MCubeConfig()
        self._g = None
        self._iso = 0.0
        self._triangles = 0

    def setVolumeTexture(self, gfusion):
        """The GFusion whose volume is meshed (getVolume() in the reference)."""
        self._g = gfusion

    def setConfig(self, cfg):
        self._cfg = cfg

    def init(self):
        """Shader build + table upload in the reference; the tables live in the
        library and are uploaded on first use."""

    def setIsolevel(self, iso):
        self._iso = float(iso)

    def generateMarchingCubes(self):
        n = C.c_uint64()
        K.check(K.lib().kf_mesh_generate(self._g.session, C.c_float(self._iso), C.byref(n)))
        self._triangles = n.value
        return self._triangles

    def getNumberTriangles(self):
        """mcubes.h:76-79 returns m_totalSum, which holds the vertex count
        (three per triangle)."""
        return 3 * self._triangles

    def triangles(self):
        return self._triangles

    def getPosBuffer(self):
        """The vertex buffer as (triangles, 3, 4) float32: (x, y, z, 0) in
        voxel-index units."""
        out = np.zeros((max(self._triangles, 1), 3, 4), np.float32)
        K.check(K.lib().kf_mesh_read(self._g.session, out.ctypes.data_as(C.POINTER(C.c_float)),
                                     self._triangles))
        return out[:self._triangles]

    def exportMesh(self, path="data/meshes/marchingCubesBin.stl"):
        K.check(K.lib().kf_mesh_export_stl(self._g.session, path.encode()))

    def exportPointCloud(self, path="data/meshes/marchingCubesBin.ply"):
        K.check(K.lib().kf_mesh_export_ply(self._g.session, path.encode()))


def mc_tables():
    """The library's triangle table (tri[256, 16], ntri[256]); host code only."""
    tri = np.zeros((256, 16), np.uint8)
    nt = np.zeros(256, np.uint8)
    K.check(K.lib().kf_mc_tables(tri.ctypes.data, nt.ctypes.data))
    return tri, nt
