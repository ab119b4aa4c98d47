"""vecruntime_amd -- MI355X-native hash aggregation and hash join for VecRuntime's kernels path.

The device library is vecruntime_amd/libvrgpu.so (C ABI: include/vr.h); `vecruntime_amd.vr`
is its Python front end.
"""
