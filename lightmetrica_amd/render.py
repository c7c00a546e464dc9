"""Command-line renderer over the GPU path, with the reference CLI's arguments
(src/lightmetrica/main.cpp:168-180):

    python -m lightmetrica_amd.render scene.xml [out.hdr] [--num-samples N] [--device D]
                                      [--film-precision fp32|fp32c|fp64]

Loads a <nanon> scene file (lightmetrica_amd.scene_xml), renders it with the file's
renderer (pt, pt.direct, pt.mis or raycast; the `.hip` spellings of the plugin are
accepted) on the GPU, and saves the film as Radiance .hdr (the reference's default output,
`result.hdr` when no path is given) or .pfm.  Every render runs on the HIP kernels: the
command fails when liblmgpu.so or the GPU is missing.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m lightmetrica_amd.render")
    ap.add_argument("config", help="scene file (<nanon version=\"1.0.dev\">)")
    ap.add_argument("output_image", nargs="?", default="", help="output image (.hdr or .pfm); default result.hdr")
    ap.add_argument("-n", "--num-samples", type=int, default=None, help="override the scheduler's num_samples")
    ap.add_argument("-d", "--device", type=int, default=0)
    ap.add_argument("--film-precision", choices=["fp32", "fp32c", "fp64"], default="fp32")
    args = ap.parse_args(argv)

    from . import film_io
    from .gpu import Device
    from .scene_xml import load_scene_xml

    f = load_scene_xml(Path(args.config), num_samples=args.num_samples)
    out = Path(args.output_image or "result.hdr")
    if out.suffix not in (".hdr", ".pfm"):
        print(f"Invalid extension '{out.suffix}'. Using 'result.hdr' as default.", file=sys.stderr)
        out = Path("result.hdr")
    t0 = time.time()
    with Device(args.device) as dev:
        dev.upload(f.scene)
        info = dev.build_bvh()
        if f.renderer == "raycast":
            img = dev.raycast(out_precision="fp64")
            what = f"raycast {f.scene.camera.width}x{f.scene.camera.height}"
        else:
            img, st = dev.render(f.renderer, film_precision=args.film_precision, out_precision="fp64")
            what = (f"{f.renderer}: {st.samples} samples, {st.closest_rays + st.shadow_rays} rays, "
                    f"{st.seconds:.3f} s on the GPU ({st.samples / max(st.seconds, 1e-12) / 1e6:.1f} M samples/s)")
    (film_io.write_pfm if out.suffix == ".pfm" else film_io.write_hdr)(out, img)
    print(f"{f.scene.num_triangles} triangles, BVH {info.num_nodes} nodes in {info.build_seconds:.2f} s; {what}; "
          f"saved {out} ({time.time() - t0:.1f} s total)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
