"""The ablation pipeline file (--pipelines): parse_pipeline_set
(src/app/pipeline_config.cpp:51-104,251-279).

    reference = "baseline"            # optional; empty selects the first
    compare-every-n-frames = 5        # optional; <= 0 disables comparisons

    [[pipeline]]
    name = "baseline"

    [[pipeline]]
    name = "bf16"
    voxel = "bf16"

Each [[pipeline]] starts from the CLI's pipeline and overrides the keys it
names: name, space, voxel, color, storage, sparse-capacity, integration,
raycast, tsdf-variant, projective-distance, distance-scaled-truncation,
truncation-distance-scale, observation-weight, max-weight, icp-damping,
icp-lambda, icp-adaptive-damping.  The volume geometry is not overridable.
Unknown keys, wrong value types, a missing name or a name with a comma raise
ValueError.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field
from pathlib import Path

from ..pipeline import PipelineConfig
from . import names


@dataclass
class PipelineSetConfig:
    """PipelineSetConfig (include/kinectfusion/pipeline_set.hpp:18-22), plus
    the per-pipeline ICP settings the reference keeps in PipelineConfig."""
    pipelines: list[PipelineConfig] = field(default_factory=list)
    reference: str = ""
    compare_every_n_frames: int = 1
    icp_damping: dict[str, tuple[str, float, bool]] = field(default_factory=dict)


def _value(node, key: str, kinds):
    # TOML booleans are not integers here, as in toml++'s value<int64_t>()
    if isinstance(node, bool) and bool not in kinds:
        raise ValueError(f"Pipeline config key '{key}' has the wrong type")
    if not isinstance(node, kinds):
        raise ValueError(f"Pipeline config key '{key}' has the wrong type")
    return node


def _float(node, key: str) -> float:
    return float(_value(node, key, (float, int)))


def parse_pipeline(table: dict, defaults: PipelineConfig, icp_defaults: tuple[str, float, bool]):
    """One [[pipeline]] table: (PipelineConfig, (damping, lambda, adaptive))."""
    config = copy.deepcopy(defaults)
    config.name = ""
    damping, lam, adaptive = icp_defaults
    for key, node in table.items():
        if key == "name":
            config.name = _value(node, key, (str,))
        elif key == "space":
            config.space = names.memory_space_from_name(_value(node, key, (str,)))
        elif key == "voxel":
            config.voxel = names.voxel_store_from_name(_value(node, key, (str,)))
        elif key == "color":
            config.color = names.color_store_from_name(_value(node, key, (str,)))
        elif key == "storage":
            config.storage = names.storage_layout_from_name(_value(node, key, (str,)))
        elif key == "sparse-capacity":
            config.block_capacity = int(_value(node, key, (int,)))
        elif key == "integration":
            config.integration.mode = names.integration_mode_from_name(_value(node, key, (str,)))
        elif key == "raycast":
            config.raycast_backend = names.raycast_backend_from_name(_value(node, key, (str,)))
        elif key == "tsdf-variant":
            config.tsdf_rule = names.tsdf_rule_from_name(_value(node, key, (str,)))
        elif key == "projective-distance":
            config.integration.projective_distance = _value(node, key, (bool,))
        elif key == "distance-scaled-truncation":
            config.integration.distance_scaled_truncation = _value(node, key, (bool,))
        elif key == "truncation-distance-scale":
            config.integration.truncation_distance_scale = _float(node, key)
        elif key == "observation-weight":
            config.integration.observation_weight = _float(node, key)
        elif key == "max-weight":
            config.integration.max_weight = _float(node, key)
        elif key == "icp-damping":
            damping = names.icp_damping_mode_from_name(_value(node, key, (str,)))
        elif key == "icp-lambda":
            lam = _float(node, key)
        elif key == "icp-adaptive-damping":
            adaptive = _value(node, key, (bool,))
        else:
            raise ValueError(f"Unknown pipeline config key '{key}'")
    if not config.name:
        raise ValueError("Every [[pipeline]] requires a name")
    if "," in config.name:
        raise ValueError("Pipeline names must not contain commas")
    return config, (damping, lam, adaptive)


def parse_pipeline_set(path, defaults: PipelineConfig, default_compare_every: int,
                       icp_defaults: tuple[str, float, bool] = ("none", 1.0e-3, False)) -> PipelineSetConfig:
    import tomli

    try:
        table = tomli.loads(Path(path).read_text())
    except (OSError, tomli.TOMLDecodeError) as err:
        raise ValueError(f"Failed to parse pipeline config {path}: {err}") from err
    out = PipelineSetConfig(compare_every_n_frames=default_compare_every)
    for key, node in table.items():
        if key == "reference":
            out.reference = _value(node, key, (str,))
        elif key == "compare-every-n-frames":
            out.compare_every_n_frames = int(_value(node, key, (int,)))
        elif key == "pipeline":
            if not isinstance(node, list) or not all(isinstance(e, dict) for e in node):
                raise ValueError("'pipeline' must be an array of tables ([[pipeline]])")
            for entry in node:
                cfg, icp = parse_pipeline(entry, defaults, icp_defaults)
                out.pipelines.append(cfg)
                out.icp_damping[cfg.name] = icp
        else:
            raise ValueError(f"Unknown pipeline config key '{key}'")
    if not out.pipelines:
        raise ValueError("Pipeline config requires at least one [[pipeline]]")
    return out
