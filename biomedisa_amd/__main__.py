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
    p.add_argument("path_to_data", type=str, metavar="PATH_TO_IMAGE", help="Location of image data")
    p.add_argument("path_to_labels", type=str, metavar="PATH_TO_LABELS", help="Location of label data")
    p.add_argument("--nbrw", type=int, default=10, help="Number of random walks starting at each pre-segmented pixel")
    p.add_argument("--sorw", type=int, default=4000, help="Steps of a random walk")
    p.add_argument("--acwe", action="store_true", default=False, help="Post-processing with active contour")
    p.add_argument("--acwe_alpha", metavar="ALPHA", type=float, default=1.0, help="Pushing force of active contour")
    p.add_argument("--acwe_smooth", metavar="SMOOTH", type=int, default=1,
                   help="Smoothing steps of active contour")
    p.add_argument("--acwe_steps", metavar="STEPS", type=int, default=3, help="Iterations of active contour")
    p.add_argument("-nc", "--no-compression", dest="compression", action="store_false",
                   help="Disable compression of segmentation results")
    p.add_argument("-allx", "--allaxis", action="store_true", default=False,
                   help="If pre-segmentation is not exlusively in xy-plane")
    p.add_argument("-d", "--denoise", action="store_true", default=False,
                   help="Smooth/denoise image data before processing")
    p.add_argument("-u", "--uncertainty", action="store_true", default=False,
                   help="Return uncertainty of segmentation result")
    p.add_argument("-i", "--ignore", type=str, default="none", help="Ignore specific label(s), e.g. 2,5,6")
    p.add_argument("-o", "--only", type=str, default="all", help="Segment only specific label(s), e.g. 1,3,5")
    p.add_argument("-s", "--smooth", nargs="?", type=int, const=100, default=0,
                   help="Number of smoothing iterations for segmentation result")
    p.add_argument("-c", "--clean", nargs="?", type=float, const=0.1, default=None,
                   help="Remove outliers, e.g. 0.5 means that objects smaller than 50 percent of the size of "
                        "the largest object will be removed")
    p.add_argument("-f", "--fill", nargs="?", type=float, const=0.9, default=None,
                   help="Fill holes, e.g. 0.5 means that all holes smaller than 50 percent of the entire label "
                        "will be filled")
    p.add_argument("-p", "--platform", default=None, help="Accepted for compatibility; the engine runs on HIP")
    p.add_argument("-rh", "--return_hits", action="store_true", default=False,
                   help="Return hits from each label")
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
