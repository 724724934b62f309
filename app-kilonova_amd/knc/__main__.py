"""``python -m knc <input.yaml|chkpt.cbor|preset> [overrides ...]``: the
reference's command line (src/main.rs:104-151) on the GPU."""
import sys

from .app import main

if __name__ == "__main__":
    sys.exit(main())
