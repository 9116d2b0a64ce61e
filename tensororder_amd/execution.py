"""The execution driver (``src/execution.py:122-152``).

``run`` contracts a sliced plan with a tensor library; when the library raises
``OutOfMemoryError`` it slices once more (``slicer.slice_once``) and restarts every
slice.  Timeouts, host memory errors and any other exception are reported as an
``Error`` line and give ``None``.
"""

import sys
import traceback

from .api import OutOfMemoryError


def run(plan, tensor_library, slicer, slice_cutoff, log=None):
    log = log or (lambda key, value: print("%s: %s" % (key, value), flush=True))
    result = None
    try:
        while True:
            try:
                result = tensor_library.contract_sliced(plan, num_slice_limit=slice_cutoff)
                break
            except OutOfMemoryError:
                slicer.slice_once(plan)
    except TimeoutError:
        log("Error", "Timeout during execution")
    except MemoryError:
        log("Error", "Out of Memory during execution")
    except Exception:
        sys.stderr.write(traceback.format_exc())
        log("Error", "Exception during execution")
    return result
