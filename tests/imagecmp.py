"""Image comparison as the reference's bench harness does it
(src/bench/image/image_io.cpp:16-20, 223-238), re-exported from the package."""
from umbreon_amd.image_io import linear_to_srgb, psnr_srgb  # noqa: F401
