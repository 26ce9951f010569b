"""The reconstruction app over the HIP pipeline (the reference's src/app):
command-line options and the --pipelines ablation file, TUM RGB-D and
synthetic frame sources, the pipeline set with its comparator, raycast PNG /
point-cloud PLY / mesh PLY / ablation CSV output, trajectories and ATE.

    python -m kinectfusion_amd.app <dataset> [options]
"""
