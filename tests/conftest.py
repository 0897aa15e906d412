import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    """Generated benchmark scenes (reference scene-JSON + OBJ format)."""
    from cudaraytracer_amd import scenes

    d = tmp_path_factory.mktemp("scenes")
    paths = {}
    for name, fn in scenes.WORKLOADS.items():
        paths[name] = scenes.write_scene(fn(), str(d / name), name)
    # Reduced-resolution copy of the spheres scene for full-image parity runs.
    paths["cornell_spheres_128"] = scenes.write_scene(scenes.cornell_spheres(width=128, height=128, spp=4),
                                                      str(d / "cornell_spheres_128"), "cornell_spheres_128")
    # Reduced terrain + subsurface sphere + HDR environment light.
    paths["terrain_small"] = scenes.write_scene(
        scenes.terrain_env(grid=101, env_w=256, env_h=128, width=160, height=90, spp=4),
        str(d / "terrain_small"), "terrain_small")
    # The same terrain with the reference's Burley subsurface profile.
    paths["terrain_small_burley"] = scenes.write_scene(
        scenes.terrain_env(grid=101, env_w=256, env_h=128, width=160, height=90, spp=4, sss="burley"),
        str(d / "terrain_small_burley"), "terrain_small_burley")
    # Small spheres scene for the real-time frame loop (G-buffer + SVGF).
    paths["cornell_spheres_rt"] = scenes.write_scene(scenes.cornell_spheres(width=96, height=64, spp=1, depth=5),
                                                     str(d / "cornell_spheres_rt"), "cornell_spheres_rt")
    return paths


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import rtref

    rtref.build()
    return rtref


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)

        return torch.cuda.device_count() > 0
    except Exception:
        return False
