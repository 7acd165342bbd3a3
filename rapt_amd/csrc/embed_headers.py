"""Build step of librapt_hip.so: writes the headers that the run-time
compiled tracer kernels include (user-written fields, rapt_jit.hip) into one
C++ include file, as raw string literals under their include names, so the
library compiles those kernels without its source tree.

usage: embed_headers.py OUT.inc HEADER...
"""
import os
import sys

DELIM = "rapt_hdr_"


def main(argv):
    out, headers = argv[1], argv[2:]
    names, texts = [], []
    for path in headers:
        text = open(path).read()
        if (")" + DELIM + '"') in text:
            raise SystemExit("%s contains the raw-string delimiter" % path)
        names.append(os.path.basename(path))
        texts.append(text)
    lines = ["// written by embed_headers.py from %s; not edited by hand" % ", ".join(names),
             "static const int kJitHeaderCount = %d;" % len(names),
             "static const char* const kJitHeaderNames[] = {"]
    lines += ['    "%s",' % n for n in names]
    lines += ["};", "static const char* const kJitHeaderTexts[] = {"]
    for t in texts:
        lines.append('R"%s(%s)%s",' % (DELIM, t, DELIM))
    lines.append("};")
    tmp = out + ".tmp"
    with open(tmp, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, out)


if __name__ == "__main__":
    main(sys.argv)
