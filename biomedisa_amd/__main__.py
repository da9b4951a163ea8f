"""Command line of the drop-in, as biomedisa's (biomedisa/interpolation.py:286-348):

    python -m biomedisa_amd PATH_TO_IMAGE PATH_TO_LABELS [options]

The results are written next to the image as final.<image name> with the label
file's type, plus .smooth / .cleaned / .filled / .cleaned.filled / .acwe /
.refined / .uncertainty files when requested (biomedisa_amd.io lists the file
types read and written).  Options of the reference that select its platform,
its web service or 3D Slicer are accepted and have no effect here.
"""
import argparse
import sys


def parser():
    p = argparse.ArgumentParser(prog="python -m biomedisa_amd", description="Biomedisa interpolation on one MI355X.",
                                formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("path_to_data", type=str, metavar="PATH_TO_IMAGE", help="image volume file (or folder of slices)")
    p.add_argument("path_to_labels", type=str, metavar="PATH_TO_LABELS", help="label volume with the pre-segmented slices")
    p.add_argument("--nbrw", type=int, default=10, help="walks started from every seed voxel")
    p.add_argument("--sorw", type=int, default=4000, help="steps per walk")
    p.add_argument("--acwe", action="store_true", default=False, help="also refine the result with an active-contour (Chan-Vese) pass")
    p.add_argument("--acwe_alpha", metavar="ALPHA", type=float, default=1.0, help="active-contour balloon force")
    p.add_argument("--acwe_smooth", metavar="SMOOTH", type=int, default=1,
                   help="curvature steps per active-contour iteration")
    p.add_argument("--acwe_steps", metavar="STEPS", type=int, default=3, help="active-contour iterations")
    p.add_argument("-nc", "--no-compression", dest="compression", action="store_false",
                   help="write result files uncompressed")
    p.add_argument("-allx", "--allaxis", action="store_true", default=False,
                   help="seeds come from labelled planes of all three axes, not only z-planes")
    p.add_argument("-d", "--denoise", action="store_true", default=False,
                   help="apply the 3x3x3 denoising filter to the image first")
    p.add_argument("-u", "--uncertainty", action="store_true", default=False,
                   help="also write the per-voxel uncertainty volume")
    p.add_argument("-i", "--ignore", type=str, default="none", help="comma-separated label values to leave out, or 'none'")
    p.add_argument("-o", "--only", type=str, default="all", help="comma-separated label values to keep, or 'all'")
    p.add_argument("-s", "--smooth", nargs="?", type=int, const=100, default=0,
                   help="curvature-smoothing iterations of the result (flag alone: 100)")
    p.add_argument("-c", "--clean", nargs="?", type=float, const=0.1, default=None,
                   help="drop connected pieces of a label below this fraction of its largest piece "
                        "(flag alone: 0.1)")
    p.add_argument("-f", "--fill", nargs="?", type=float, const=0.9, default=None,
                   help="fill enclosed holes smaller than this fraction of the label's size (flag alone: 0.9)")
    p.add_argument("-p", "--platform", default=None, help="Accepted for compatibility; the engine runs on HIP")
    p.add_argument("-rh", "--return_hits", action="store_true", default=False,
                   help="also return the per-label hit volumes")
    p.add_argument("-iid", "--img_id", type=str, default=None, help="Accepted for compatibility")
    p.add_argument("-lid", "--label_id", type=str, default=None, help="Accepted for compatibility")
    p.add_argument("-r", "--remote", action="store_true", default=False, help="Accepted for compatibility")
    p.add_argument("-q", "--queue", type=int, default=0, help="Accepted for compatibility")
    p.add_argument("--slicer", action="store_true", default=False, help="Accepted for compatibility")
    # this engine's own options
    p.add_argument("--method", choices=["plain", "guided"], default="plain",
                   help="plain walker, or the surrogate-guided pipeline")
    p.add_argument("--path", choices=["auto", "small", "large"], default="auto",
                   help="rules of the reference's small- or large-volume driver ('auto': by size)")
    p.add_argument("--device", type=int, default=0, help="HIP device ordinal")
    return p


def main(argv=None):
    kwargs = vars(parser().parse_args(argv))
    from .interpolation import smart_interpolation
    res = smart_interpolation(None, None, **kwargs)
    return 0 if res is not None else 1


if __name__ == "__main__":
    sys.exit(main())
