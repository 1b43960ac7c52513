"""Copies the reference's scene assets (data files: the scene JSONs and the OBJ
boundary mesh of FluidSimulationProject/assets/simulationScenes/, and the
simulation profiles of FluidSimulationProject/assets/simulationProfiles/) into
tests/golden/app/assets/, keeping the application's layout so that the meshes'
relative paths resolve, for tests that run where the reference tree is absent.

    python tests/golden/copy_scene_assets.py [reference root, default /root/reference]
"""
import os
import shutil
import sys

FILES = ["triangleTestScene.json", "untitled.obj", "untitled.mtl", "scene.json", "debugSmallestScene.json"]
PROFILES = ["increasingBoxSizeProfiles.json", "increasingReorderTimeIntervalProfiles.json",
            "increasingStaticParticleCountProfiles.json", "systemProfilingProfiles.json",
            "systemVisualisationProfile.json", "waveDetailedProfilingProfile.json",
            "waveIncreasingParticleReorderTimeIntervalDetailedProfilingProfiles.json", "waveProfile.json"]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "FluidSimulationProject", "assets", "simulationScenes")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "app", "assets", "simulationScenes")
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        print(f"copied {f}")
    src = os.path.join(ref, "FluidSimulationProject", "assets", "simulationProfiles")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "app", "assets", "simulationProfiles")
    os.makedirs(dst, exist_ok=True)
    for f in PROFILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        print(f"copied {f}")


if __name__ == "__main__":
    main()
