"""umapers_amd — MI355X-native UMAP fit and transform (HIP kernels for gfx950).

The public surface follows the umapers Python facade
(umap_rs/python/umapers/_api.py): ``Umap(...).fit``, ``fit_transform``,
``fit_transform_with_knn``, ``transform`` and ``profile_fit_transform``.
"""
from ._api import Umap, fit_transform

__all__ = ["Umap", "fit_transform"]
