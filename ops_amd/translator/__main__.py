import argparse
import sys

from .hip_gfx950 import TranslateError, translate_app


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="ops-translator")
    ap.add_argument("-t", "--target", default="hip_gfx950", choices=["hip_gfx950"])
    ap.add_argument("-o", "--out", required=True, help="output directory")
    ap.add_argument("--app", required=True, help="application name (names the kernels TU)")
    ap.add_argument("--file_paths", nargs="+", required=True, help="OPS sources")
    a = ap.parse_args(argv)
    try:
        r = translate_app(a.app, a.file_paths, a.out)
    except TranslateError as e:
        print(f"ops-translator: error: {e}", file=sys.stderr)
        return 1
    print(r["kernels"])
    for h in r["host"]:
        print(h)
    return 0


if __name__ == "__main__":
    sys.exit(main())
