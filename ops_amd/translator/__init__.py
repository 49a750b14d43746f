"""OPS source-to-source translator with a single target, `hip_gfx950`.

Command line (mirrors the reference's `ops-translator -t <target> --file_paths ...`,
ops_translator/ops-translator/__main__.py:22-46):

    python -m ops_amd.translator -t hip_gfx950 -o OUT_DIR --app NAME --file_paths a.cpp b.cpp
"""
from .hip_gfx950 import TranslateError, translate_app  # noqa: F401
