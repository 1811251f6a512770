"""MI355X-native shallow-water step backend (HIP kernels behind a ctypes C ABI).

``inundation_amd.workflow`` mirrors the reference's ``SWEWorkflow`` API with one
more backend, ``"hip_resident"``; ``inundation_amd.hip`` holds that backend.
"""
