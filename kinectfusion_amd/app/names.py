"""Name <-> value mappings shared by the CLI and the TOML pipeline file
(src/app/pipeline_config.cpp:113-249).  Unknown names raise ValueError (the
reference's std::invalid_argument) with the reference's messages; values
are the names PipelineConfig uses."""
from __future__ import annotations


def _pick(kind: str, name: str, table: dict[str, str], expected: str) -> str:
    if name in table:
        return table[name]
    raise ValueError(f"Unknown {kind} '{name}' (expected {expected})")


# The memory spaces the CLI and the TOML file name.  "hip" is this build's
# space; "cpu" and "cuda" are the reference's and are refused when a
# pipeline is created (they are not built here).
MEMORY_SPACES = {"cpu": "cpu", "cuda": "cuda", "hip": "hip"}


def memory_space_from_name(name: str) -> str:
    return _pick("memory space", name.lower(), MEMORY_SPACES, "'cpu', 'cuda' or 'hip'")


def tsdf_rule_from_name(name: str) -> str:
    return _pick("TSDF variant", name, {"classic": "classic", "angle-weighted": "angle-weighted"},
                 "'classic' or 'angle-weighted'")


def voxel_store_from_name(name: str) -> str:
    return _pick("voxel store", name, {"float": "float", "quantized": "quantized", "bf16": "bf16"},
                 "'float', 'quantized', or 'bf16'")


def color_store_from_name(name: str) -> str:
    return _pick("color store", name, {"float": "float", "none": "none"}, "'float' or 'none'")


def storage_layout_from_name(name: str) -> str:
    """"sparse" is the hashed block store of this build."""
    return _pick("storage layout", name, {"dense": "dense", "sparse": "hashed"}, "'dense' or 'sparse'")


def icp_damping_mode_from_name(name: str) -> str:
    return _pick("ICP damping mode", name, {"none": "none", "identity": "identity", "diagonal": "diagonal"},
                 "'none', 'identity', or 'diagonal'")


def integration_mode_from_name(name: str) -> str:
    return _pick("integration mode", name, {"full": "full", "band": "band"}, "'full' or 'band'")


def raycast_backend_from_name(name: str) -> str:
    return _pick("raycast backend", name, {"march": "march", "bitmap-march": "bitmap-march",
                                            "band-march": "band-march"},
                 "'march', 'bitmap-march', or 'band-march'")
