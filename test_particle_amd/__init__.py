"""MI355X test-particle tracker: HIP kernels behind a C ABI (libtpgpu.so)."""
