"""``python -m dftpy_amd config.ini``: the reference's ``dftpy`` command
(cui/main.py) for density-optimization jobs on the GPU path."""
from .interface import main

if __name__ == "__main__":
    main()
