"""holodeck_amd: MI355X-native (gfx950) semi-analytic GW-background hot path of holodeck.

Drop-in surface for holodeck's SAM GWB path: ``sams.Semi_Analytic_Model``
(``static_binary_density``, ``dynamic_binary_number_at_fobs``, ``gwb``),
``hardening.Hard_GW`` / ``Fixed_Time_2PL_SAM``, ``host_relations.MMBulge_*``,
``sams.sam_cyutils`` and ``cyutils`` (same names and signatures as the
reference's Cython modules), ``single_sources.ss_gws_redz``.  All per-cell
compute runs in hand-written HIP kernels in ``libholohip.so``.
"""
__version__ = "0.1.0"

from . import constants  # noqa
from . import utils  # noqa
from .cosmology import cosmo, Cosmology  # noqa
from . import host_relations  # noqa
from . import hardening  # noqa
from . import sams  # noqa
from . import cyutils  # noqa
from . import single_sources  # noqa
from . import gravwaves  # noqa

__all__ = ["cosmo", "Cosmology", "sams", "hardening", "host_relations", "cyutils", "single_sources", "gravwaves",
           "utils", "constants"]
