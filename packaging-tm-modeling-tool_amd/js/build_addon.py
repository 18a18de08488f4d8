"""Build the Node-API addon fcbga_hip.node next to libfcbga_hip.so (g++, Node headers)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
NODE_INC = "/usr/include/node"


def main():
    if not os.path.exists(os.path.join(NODE_INC, "node_api.h")):
        print("node headers not found: addon not built", file=sys.stderr)
        return
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + NODE_INC,
           "-o", os.path.join(PKG, "fcbga_hip.node"), os.path.join(HERE, "fcbga_napi.cc"),
           "-L" + PKG, "-l:libfcbga_hip.so", "-Wl,-rpath,$ORIGIN"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    main()
