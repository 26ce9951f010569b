"""`python -m kinectfusion_amd.app`: main (src/app/main.cpp:35-68): parse
the options, validate them, run the reconstruction; any error is reported
and exits with status 1."""
from __future__ import annotations

import logging
import sys


def main(argv=None) -> int:
    from .options import parse_args, validate_options
    from .reconstruction import Reconstruction

    logging.basicConfig(level=logging.INFO, format="[%(levelname)s] %(message)s", stream=sys.stderr)
    log = logging.getLogger("kinectfusion")
    try:
        options = parse_args(argv)
        validate_options(options)
        log.info("Starting KinectFusion: dataset=%s max_frames=%d volume=%d^3 voxel_size=%g truncation_distance=%g",
                 options.dataset_dir, options.max_frames, options.volume_resolution, options.voxel_size,
                 options.truncation_distance)
        return Reconstruction(options).run()
    except Exception as err:  # noqa: BLE001 - the reference's catch-all in main
        log.error("Unhandled exception in main: %s", err)
        return 1


if __name__ == "__main__":
    sys.exit(main())
