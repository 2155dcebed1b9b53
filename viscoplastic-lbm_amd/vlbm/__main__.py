"""Runs one simulation from a pyvisc-style YAML file on the GPU:

    python -m vlbm -c conf.yaml

(same -c/--conf-path argument as pyvisc/args_handler.py:10-33; the first
entry of the YAML list is simulated, writing PATH/ID/ID_*.bin and ID_info.txt).
"""
import argparse
import sys

from .config import read_configs
from .runner import run


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="vlbm")
    ap.add_argument("-c", "--conf-path", required=True, help="Path to the YAML configuration")
    args = ap.parse_args(argv)
    cfg = read_configs(args.conf_path)
    res = run(cfg)
    print(f"steps {res.steps}  time {res.seconds:.3f} s  MLUPS {res.mlups:.1f}"
          + (f"  stopped: {res.stopped}" if res.stopped else ""))
    return 0


if __name__ == "__main__":
    sys.exit(main())
