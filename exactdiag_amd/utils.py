"""Stage-level memoisation over files on disk (reference: utils.py:9-25)."""

from __future__ import annotations

import logging
import pathlib
from os.path import isfile

_logger = logging.getLogger(__name__)


def load_calculate_save(*paths: pathlib.Path, load_fun, calc_fun, save_fun):
    """Load all paths if they all exist, else calculate and save (utils.py:9-25).

    An existing file that fails to load raises OSError, as in the reference.
    """
    already_exist = all(isfile(path) for path in paths)
    str_paths = [str(p) for p in paths]
    if already_exist:
        try:
            result = load_fun(*paths)
            _logger.debug(f"loaded {str_paths}")
        except OSError as e:
            message = f"Loading {str_paths}: file(s) exists but load failed. {e}"
            _logger.debug(message)
            raise OSError(message) from e
    else:
        _logger.debug(f"starting calculating {str_paths}")
        result = calc_fun()
        save_fun(result, *paths)
        _logger.debug(f"finished calculating {str_paths}")
    return result
