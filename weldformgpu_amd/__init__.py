"""weldformgpu_amd — MI355X-native (gfx950 HIP) WeldFormGPU SPH step pipeline."""
