"""Frame files of a dataset, as Dataset::Dataset finds and reads them
(Dataset.cpp:169-227): prefix + i + suffix for i = 1, 2, ... while the file
exists (at most `count` when count >= 0), each read as BGEO positions
(Dataset::ReadFile, Dataset.cpp:292-306) by the library's pm_bgeo_read."""
from __future__ import annotations

import pathlib

from ._lib import bgeo_read


def dataset_files(prefix: str, suffix: str, count: int = -1) -> list[pathlib.Path]:
    files = []
    i = 1
    while count < 0 or i <= count:
        p = pathlib.Path(f"{prefix}{i}{suffix}")
        if not p.exists():
            break
        files.append(p)
        i += 1
    return files


def read_dataset(prefix: str, suffix: str, count: int = -1):
    """Positions (n, 3) float32 of every frame file, in frame order."""
    return [bgeo_read(p) for p in dataset_files(prefix, suffix, count)]
