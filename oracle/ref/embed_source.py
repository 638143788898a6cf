"""Write a C file that defines `const char <name>[]` holding the text of a
source file (build step of the drop-in recipes in this directory: the
generic device program compiles that text for the GPU at run time).

    python3 embed_source.py NAME FILE > out.c
"""
import sys


def main():
    name, path = sys.argv[1], sys.argv[2]
    text = open(path, "rb").read().decode("ascii")
    out = [f"const char {name}[] ="]
    for line in text.splitlines(True):
        esc = line.replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n").replace("\t", "\\t")
        out.append(f'\t"{esc}"')
    out.append("\t;")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
