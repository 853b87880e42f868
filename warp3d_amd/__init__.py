"""warp3d_amd: MI355X-native element-block hot path of WARP3D."""
