"""MI355X implicit cloth solve: host mirror of the reference's GpuSolver over libbd.so (HIP, gfx950).

Modules: scene (scene compiler), material, noise, configs (workload scenes), capi (C-ABI structs and
library loader), solver (GpuSolver).
"""
