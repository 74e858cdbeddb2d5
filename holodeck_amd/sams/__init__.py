"""Semi-analytic models (holodeck/sams)."""
from .components import (  # noqa
    GSMF_Schechter, GPF_Power_Law, GMT_Power_Law, GMR_Illustris,
    _Galaxy_Pair_Fraction, _Galaxy_Stellar_Mass_Function, _Galaxy_Merger_Time, _Galaxy_Merger_Rate,
)
from .sam import Semi_Analytic_Model  # noqa
from . import sam_cyutils  # noqa
